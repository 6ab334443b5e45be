! golden_driver.f90 -- writes golden vectors of Hochunk3d's own leaf routines.
! Built only in the build container, against the reference sources where they
! lie (see Makefile), into oracle/_ref/.  Every input is either printed with
! 17 significant digits or defined by an exactly reproducible formula, so the
! CPU restatement can be fed the same inputs in the tests.
program golden_driver
  use ecuyer_cote_mod
  use random_standard_uniform_mod
  use random
  use closest_wall
  use dust_mod
  use grid_mod
  use opacin_mod
  implicit none
  integer :: gg, k, j, i, id, jj, face, idust2, n
  real*8 :: v(8), ux, uy, uz, x, y, z, r2a, r2b, a1, b1, a2, b2, t, phi1, phi2
  real*8 :: xx(12), xs, T0, na, Td, p1, p2, p3, p4, cost, cost2
  real*8 :: c1, s1, c2, s2, cp1, sp1, cp2, sp2, cn, sn, pn, ri1
  real*8 :: xran, nu, Tk
  logical(1) :: onw
  character(len=80) :: dname(ndg)
  real*8, external :: lucydustfreq, KapP, dusttemp, dustfreq, chiR, ierfc
  real*8 :: sux, suy, suz
  integer, external :: ifix_dummy

  call set_all_seeds(12345678, 87654321)

  ! ---- 1. generator streams and segment splitting
  do gg = 1, 3, 2
    call set_current_generator(gg)
    call reinitialize_current_generator(-1)
    do k = 0, 3
      if (k > 0) call reinitialize_current_generator(1)
      do j = 1, 6
        v(j) = random_standard_uniform()
      end do
      write(*,'(A,1X,I0,1X,I0,1X,6(ES26.17E3,1X))') 'rng', gg, k, v(1:6)
    end do
  end do

  ! ---- 2. bbfreq on generator 1 segment 0
  call set_current_generator(1)
  call reinitialize_current_generator(-1)
  do j = 1, 40
    T0 = 0.05d0 * dble(j)
    nu = bbfreq(T0)
    write(*,'(A,1X,2(ES26.17E3,1X))') 'bbfreq', T0, nu
  end do

  ! ---- 3. wall roots: rays drawn from generator 5
  call set_current_generator(5)
  call reinitialize_current_generator(-1)
  do n = 1, 300
    x = 40.d0 * (random_standard_uniform() - 0.5d0)
    y = 40.d0 * (random_standard_uniform() - 0.5d0)
    z = 40.d0 * (random_standard_uniform() - 0.5d0)
    uz = 1.d0 - 2.d0 * random_standard_uniform()
    phi1 = 6.283185307179586d0 * random_standard_uniform()
    ux = sqrt(1.d0 - uz*uz) * cos(phi1)
    uy = sqrt(1.d0 - uz*uz) * sin(phi1)
    r2a = (5.d0 + 20.d0 * random_standard_uniform())**2
    r2b = r2a * (1.d0 + random_standard_uniform())
    phi1 = 6.283185307179586d0 * random_standard_uniform()
    phi2 = phi1 + 0.5d0 * random_standard_uniform()
    a1 = sin(phi1); b1 = -cos(phi1); a2 = sin(phi2); b2 = -cos(phi2)
    onw = mod(n, 3) == 0
    call reset_t()
    call radface(ux, uy, uz, x, y, z, r2a, r2b)
    call phiface(a1, b1, 0.d0, 0.d0, a2, b2, 0.d0, 0.d0, ux, uy, uz, x, y, z)
    t = -1.d0; face = -1
    if (tp1 /= huge(tp1) .and. .not.(onw .and. tp2 == huge(tp2) .and. .not. (abs(tn1) < abs(tp1)))) then
      call find_next_wall(onw, t, face)
    end if
    write(*,'(A,1X,I0,1X,12(ES26.17E3,1X),I0,1X,I0,1X,I0,1X,I0,1X,ES26.17E3,1X,ES26.17E3,1X,ES26.17E3,1X,ES26.17E3)') &
      'walls', merge(1, 0, onw), ux, uy, uz, x, y, z, r2a, r2b, a1, b1, a2, b2, in1, ip1, ip2, face, &
      tn1, tp1, tp2, t
  end do

  ! ---- 4. locate (nr.f)
  do i = 1, 12
    xx(i) = dble(i*i) * 0.25d0
  end do
  do j = -2, 150
    xs = dble(j) * 0.25d0
    call locate(xx, 12, xs, k)
    write(*,'(A,1X,ES26.17E3,1X,I0)') 'locate', xs, k
  end do

  ! ---- 5-7. dust tables from the reference's dustinit/setemiss on a
  !      synthetic dust file written by the Makefile (synthetic_dust.par)
  dname = 'synthetic_dust.par'
  call dustinit(dname)
  write(*,'(A,1X,I0,1X,ES26.17E3)') 'kappav', nlambda(1), kappav(1)
  do i = 1, nlambda(1)
    write(*,'(A,1X,I0,1X,4(ES26.17E3,1X))') 'dustarr', i, nudust(i,1), kappad(i,1), adust(i,1), gdust(i,1)
  end do
  do j = 1, 300
    nu = 10.d0**(-3.5d0 + 6.5d0 * dble(j-1) / 299.d0)
    call opacset(nu)
    write(*,'(A,1X,5(ES26.17E3,1X))') 'opacset', nu, kappa(1), albedo(1), g(1), pl(1)
  end do
  do i = 1, nT, 37
    do jj = 1, nnu, 31
      write(*,'(A,1X,I0,1X,I0,1X,ES26.17E3)') 'kapBint', i, jj, kapBint(i,jj,1)
    end do
  end do
  do i = 1, nT
    write(*,'(A,1X,I0,1X,2(ES26.17E3,1X))') 'kappap', i, kappap(i,1), lTint(i)
  end do
  ! the Bjorkman-Wood table of the same setemiss call (kappa_abs dB/dT)
  do i = 1, nT, 37
    do jj = 1, nnu, 31
      write(*,'(A,1X,I0,1X,I0,1X,ES26.17E3)') 'kapdBint', i, jj, kapdBint(i,jj,1)
    end do
  end do
  do j = 1, 80
    Tk = 0.12d0 * 1.1d0**j
    nu = KapP(Tk/11605.d0, 1)
    write(*,'(A,1X,2(ES26.17E3,1X))') 'kapp', Tk/11605.d0, nu
  end do
  ! the Rosseland-mean table of the same setemiss call, and chiR on it
  ! (inside the table and beyond both of its ends)
  do i = 1, nT
    write(*,'(A,1X,I0,1X,ES26.17E3)') 'kappar', i, kappar(i,1)
  end do
  do j = 1, 110
    Tk = 0.06d0 * 1.1d0**j
    nu = chiR(Tk/11605.d0, 1)
    write(*,'(A,1X,2(ES26.17E3,1X))') 'chir', Tk/11605.d0, nu
  end do

  ! ---- 7b. ierfc (mathsubs.f90) and isotrp on generator 8
  do j = 1, 60
    xs = 1.9d0 * dble(j) / 60.d0
    nu = ierfc(xs)
    write(*,'(A,1X,2(ES26.17E3,1X))') 'ierfc', xs, nu
  end do
  call set_current_generator(8)
  call reinitialize_current_generator(-1)
  do j = 1, 300
    suz = cos(0.37d0 * dble(j))
    sux = sqrt(1.d0 - suz*suz) * cos(1.3d0 * dble(j))
    suy = sqrt(1.d0 - suz*suz) * sin(1.3d0 * dble(j))
    call isotrp(sux, suy, suz, ux, uy, uz)
    write(*,'(A,1X,6(ES26.17E3,1X))') 'isotrp', sux, suy, suz, ux, uy, uz
  end do

  ! ---- 8. lucydustfreq on a formula table (exact in IEEE arithmetic)
  do id = 1, ndg
    do jj = 1, nnu
      do i = 1, nT
        kapBint(i,jj,id) = (dble(jj-1)/1024.d0)**2 * (dble(i)/1000.d0) + &
             (dble(jj-1)/1024.d0) * (1.d0 - dble(i)/1000.d0)
      end do
    end do
  end do
  call set_current_generator(2)
  call reinitialize_current_generator(-1)
  do j = 1, 200
    Tk = 0.2d0 * 1.045d0**j
    nu = lucydustfreq(Tk/11605.d0, 1)
    write(*,'(A,1X,2(ES26.17E3,1X))') 'lucy', Tk/11605.d0, nu
  end do

  ! ---- 8b. dustfreq (Bjorkman-Wood re-emission) on the same formula table
  !      in kapdBint, below and above the sublimation temperature (bbfreq)
  do id = 1, ndg
    do jj = 1, nnu
      do i = 1, nT
        kapdBint(i,jj,id) = (dble(jj-1)/1024.d0)**2 * (dble(i)/1000.d0) + &
             (dble(jj-1)/1024.d0) * (1.d0 - dble(i)/1000.d0)
      end do
    end do
  end do
  Tsub = 1600.d0
  call set_current_generator(7)
  call reinitialize_current_generator(-1)
  do j = 1, 200
    Tk = 0.2d0 * 1.046d0**j
    nu = dustfreq(Tk/11605.d0, 1)
    write(*,'(A,1X,2(ES26.17E3,1X))') 'dustfreq', Tk/11605.d0, nu
  end do

  ! ---- 9. dusttemp on a formula Planck-mean table
  do id = 1, ndg
    do i = 1, nT
      kappap(i,id) = 1.d-4 * dble(i) + 1.d-6 * dble(i*i)
    end do
  end do
  nrg = 2; ntg = 2; npg = 2
  call grid_allocate()
  ! inputs chosen where the fixed-point iteration of dusttemp converges
  do j = 1, 36
    massarr(1,1,1,1) = 1.d5 * 1.3d0**j
    na = 0.37d0 * 1.7d0**(j/2)
    T0 = dble(mod(j,5)) * 3.d0 / 11605.d0
    Td = dusttemp(1, 1, 1, na, T0, 3.839d33, 100000, 1)
    write(*,'(A,1X,4(ES26.17E3,1X))') 'dusttemp', massarr(1,1,1,1), na, T0, Td
  end do

  ! ---- 10. select_dust on generator 4
  call set_current_generator(4)
  call reinitialize_current_generator(-1)
  do id = 1, ndg
    kapd(id) = 1.d0 + 0.5d0 * dble(id)
    densarr(1,1,1,id) = merge(0.d0, 1.d0 / dble(id), mod(id,3) == 0)
  end do
  do j = 1, 200
    call select_dust(1, 1, 1, is_any, idust2)
    write(*,'(A,1X,I0)') 'select', idust2
  end do

  ! ---- 11. dustmat and findangle3 on generator 6 inputs
  call set_current_generator(6)
  call reinitialize_current_generator(-1)
  do j = 1, 200
    g(1) = 0.9d0 * (random_standard_uniform() - 0.3d0)
    pl(1) = 0.6d0 * random_standard_uniform()
    g2(1) = g(1)*g(1); g2p1(1) = g2(1) + 1.d0; onemg2(1) = 1.d0 - g2(1); twog(1) = 2.d0*g(1)
    cost = 1.d0 - 2.d0 * random_standard_uniform()
    cost2 = cost*cost
    call dustmat(p1, p2, p3, p4, cost, cost2, 1)
    write(*,'(A,1X,6(ES26.17E3,1X))') 'dustmat', g(1), pl(1), cost, p1, p2, p3
  end do
  do j = 1, 300
    c1 = 1.d0 - 2.d0 * random_standard_uniform(); s1 = sqrt(1.d0 - c1*c1)
    phi1 = 6.283185307179586d0 * random_standard_uniform(); cp1 = cos(phi1); sp1 = sin(phi1)
    c2 = 1.d0 - 2.d0 * random_standard_uniform(); s2 = sqrt(1.d0 - c2*c2)
    phi2 = 6.283185307179586d0 * random_standard_uniform(); cp2 = cos(phi2); sp2 = sin(phi2)
    ri1 = -7.d0
    call findangle3(c1, s1, phi1, cp1, sp1, c2, s2, phi2, cp2, sp2, cn, sn, pn, ri1)
    write(*,'(A,1X,14(ES26.17E3,1X))') 'findangle3', c1, s1, phi1, cp1, sp1, c2, s2, phi2, cp2, sp2, cn, sn, pn, ri1
  end do
end program golden_driver

integer function ifix_dummy()
  ifix_dummy = 0
end function ifix_dummy
