! A Fortran driver of one transport pass through hc_iface, as the
! reference's disk() would call the library: it reads a model's boundary
! arrays from a stream file (written by tests/test_gpu_fortran.py), creates a
! context, sets the grid, dust and sources, runs the pass and writes the SED,
! dtauabs and the scalar tallies to a second stream file.
program hc_pass_driver
  use iso_c_binding
  use hc_iface
  implicit none
  integer(c_int32_t) :: hdr(14)
  integer :: nrg, ntg, npg, ndg, nfreq, nmu, nph, nap, npeel, nxhst, nlammax, nT, nnu
  integer :: ncellw, nsed, u_in, u_out, ierr
  real(c_double), allocatable, target :: r2arr(:), tan2arr(:), thetarr(:), aarr(:), barr(:), &
       phiarr(:), densarr(:), tdust(:), nudust(:), kappad(:), adust(:), gdust(:), pldust(:), &
       kappav(:), lTint(:), nuint(:), kapBint(:), kappap(:), u(:), aperture2(:), coste(:), &
       sinte(:), cospe(:), sinpe(:), sed8(:), dtauabs(:)
  integer(c_int32_t), allocatable :: nlambda(:)
  real(c_double) :: tsub, d3(3), d16(16), d5(5), d24(24)
  integer(c_int32_t) :: i4(4), i2(2)
  integer(c_int64_t) :: np, seg0
  type(c_ptr) :: hc
  type(hc_dims) :: dims
  type(hc_sources) :: src
  type(hc_tallies) :: t
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u_in, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u_in) hdr
  nrg = hdr(1); ntg = hdr(2); npg = hdr(3); ndg = hdr(4); nfreq = hdr(5); nmu = hdr(6)
  nph = hdr(7); nap = hdr(8); npeel = hdr(9); nxhst = hdr(10); nlammax = hdr(11)
  nT = hdr(12); nnu = hdr(13)
  ncellw = nrg * ntg * npg
  allocate(r2arr(nrg), tan2arr(ntg), thetarr(ntg), aarr(npg), barr(npg), phiarr(npg))
  allocate(densarr(ncellw * ndg), tdust(ncellw * ndg))
  read(u_in) r2arr, tan2arr, thetarr, aarr, barr, phiarr, densarr, tdust, tsub
  allocate(nlambda(ndg), nudust(nlammax * ndg), kappad(nlammax * ndg), adust(nlammax * ndg))
  allocate(gdust(nlammax * ndg), pldust(nlammax * ndg), kappav(ndg), lTint(nT), nuint(nnu))
  allocate(kapBint(nT * nnu * ndg), kappap(nT * ndg))
  read(u_in) nlambda, nudust, kappad, adust, gdust, pldust, kappav, lTint, nuint, kapBint, &
       kappap
  ! hc_sources, component by component in the order of its declaration
  read(u_in) d3, i4, d16, i2, d5, d24
  allocate(u(nmu), aperture2(nap), coste(npeel), sinte(npeel), cospe(npeel), sinpe(npeel))
  read(u_in) u, aperture2, coste, sinte, cospe, sinpe
  read(u_in) np, seg0
  close(u_in)

  src%rstar = d3(1); src%tstar = d3(2); src%tshock = d3(3)
  src%limb = i4(1); src%ips = i4(2); src%ispot = i4(3); src%nspot = i4(4)
  src%fspot = d16(1); src%csprad = d16(2); src%cspot1 = d16(3); src%sspot1 = d16(4)
  src%scspot1 = d16(5); src%ssspot1 = d16(6); src%cspot2 = d16(7); src%sspot2 = d16(8)
  src%scspot2 = d16(9); src%ssspot2 = d16(10); src%accfrac = d16(11); src%accfrac2 = d16(12)
  src%numin = d16(13); src%lnurat = d16(14); src%rmaxi = d16(15); src%ltot = d16(16)
  src%iintmax = i2(1); src%pad = i2(2)
  src%u = c_loc(u); src%aperture2 = c_loc(aperture2)
  src%coste = c_loc(coste); src%sinte = c_loc(sinte)
  src%cospe = c_loc(cospe); src%sinpe = c_loc(sinpe)
  src%rddustmin = d5(1); src%rmaxdmax = d5(2); src%rmax = d5(3); src%massc = d5(4)
  src%mdot = d5(5)
  src%fmass = d24(1:8); src%z1 = d24(9:16); src%bflare = d24(17:24)

  dims = hc_dims(nrg, ntg, npg, ndg, nfreq, nmu, nph, nap, npeel, nxhst, nlammax, nT, nnu, &
                 HC_FP64, 0)
  ierr = hc_create(dims, 0_c_int, hc)
  if (ierr /= 0) stop 'hc_create'
  ierr = hc_set_grid(hc, r2arr, tan2arr, thetarr, aarr, barr, phiarr, densarr, tdust, tsub)
  if (ierr /= 0) stop 'hc_set_grid'
  ierr = hc_set_dust(hc, nlambda, nudust, kappad, adust, gdust, pldust, kappav, lTint, &
                     nuint, kapBint, kappap)
  if (ierr /= 0) stop 'hc_set_dust'
  ierr = hc_set_sources(hc, src)
  if (ierr /= 0) stop 'hc_set_sources'

  nsed = nfreq * nmu * nph * nap * 8 * 9
  allocate(sed8(nsed), dtauabs(ncellw * ndg))
  ! every output pointer is set: NULL skips that tally
  t%nabs = c_null_ptr; t%jmean = c_null_ptr; t%killed = c_null_ptr; t%dbig = c_null_ptr
  t%dsg = c_null_ptr; t%peel = c_null_ptr; t%img_i = c_null_ptr; t%img_q = c_null_ptr
  t%img_u = c_null_ptr; t%img_v = c_null_ptr; t%si_init = c_null_ptr
  t%dtauabs = c_loc(dtauabs)
  t%sed = c_loc(sed8)
  ierr = hc_run_iteration(hc, np, 0_c_int32_t, 12345678_c_int32_t, 87654321_c_int32_t, &
                          1_c_int32_t, seg0, t, c_null_ptr)
  if (ierr /= 0) stop 'hc_run_iteration'
  ierr = hc_destroy(hc)

  open(newunit=u_out, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u_out) sed8, dtauabs, t%tot, t%flux, t%nscat, t%abflux, t%sumall
  close(u_out)
  print '(A)', 'pass done'
end program hc_pass_driver
