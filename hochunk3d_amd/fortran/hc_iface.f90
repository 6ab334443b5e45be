! hc_iface.f90 -- ISO_C_BINDING interfaces of libhochunk3d_amd.so (include/hc.h)
! for a Fortran driver.  Module arrays of grid_mod / dust_mod / tts_mod are
! passed as they are: column-major, the reference's dimensions.
module hc_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: HC_FP64 = 0, HC_FP32 = 1

  type, bind(C) :: hc_dims
    integer(c_int32_t) :: nrg, ntg, npg, ndg
    integer(c_int32_t) :: nfreq, nmu, nph, nap, npeel, nxhst
    integer(c_int32_t) :: nlammax, nT, nnu
    integer(c_int32_t) :: precision, imcube
  end type hc_dims

  type, bind(C) :: hc_sources
    real(c_double) :: rstar, tstar, tshock
    integer(c_int32_t) :: limb, ips, ispot, nspot
    real(c_double) :: fspot, csprad, cspot1, sspot1, scspot1, ssspot1, cspot2, sspot2, &
                      scspot2, ssspot2
    real(c_double) :: accfrac, accfrac2, numin, lnurat, rmaxi, ltot
    integer(c_int32_t) :: iintmax, pad
    type(c_ptr) :: u, aperture2, coste, sinte, cospe, sinpe
    real(c_double) :: rddustmin, rmaxdmax, rmax
    real(c_double) :: massc, mdot
    real(c_double) :: fmass(8), z1(8), bflare(8)
  end type hc_sources

  type, bind(C) :: hc_tallies
    type(c_ptr) :: dtauabs, nabs, jmean, killed, dbig, dsg, sed, peel
    type(c_ptr) :: img_i, img_q, img_u, img_v, si_init
    real(c_double) :: tot, flux, sflux, nscat, scount, abflux, aflux, sumsub, sumall, hit, htot
    integer(c_int64_t) :: n_err_nowall, n_err_dust, n_warn
    real(c_double) :: kernel_ms
    real(c_double) :: ncross, nscat_ev, nabs_ev, npeelray, npeelcross
    real(c_double) :: max_packet_work
    real(c_double) :: ntally
  end type hc_tallies

  ! the analytic densities of grid setup (hc_build_density)
  type, bind(C) :: hc_density_model
    integer(c_int32_t) :: nrg, ntg, npg, ndisk
    type(c_ptr) :: rarr, thetarr, phiarr, fractal
    real(c_double) :: rmincgs
    integer(c_int32_t) :: disk_id(8)
    real(c_double) :: disk_fmass(8), disk_z1(8), disk_b(8)
    real(c_double) :: a_exp, rmind, rmaxd, massd
    real(c_double) :: rc, rhoe0, rmine, rmax, rhoamb
    integer(c_int32_t) :: env_id, cav_id
    real(c_double) :: c1e, cavity_exp, rho_cavity, r1000
  end type hc_density_model

  interface
    integer(c_int) function hc_create(dims, device, ctx) bind(C, name='hc_create')
      import :: c_int, c_ptr, hc_dims
      type(hc_dims), intent(in) :: dims
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
    end function hc_create

    integer(c_int) function hc_set_grid(ctx, r2arr, tan2arr, thetarr, aarr, barr, phiarr, &
                                        densarr, tdust, tsub) bind(C, name='hc_set_grid')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: r2arr(*), tan2arr(*), thetarr(*), aarr(*), barr(*), &
                                    phiarr(*), densarr(*), tdust(*)
      real(c_double), value :: tsub
    end function hc_set_grid

    integer(c_int) function hc_set_temperature(ctx, tdust) bind(C, name='hc_set_temperature')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: tdust(*)
    end function hc_set_temperature

    integer(c_int) function hc_get_temperature(ctx, tdust) bind(C, name='hc_get_temperature')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: tdust(*)
    end function hc_get_temperature

    integer(c_int) function hc_set_dust(ctx, nlambda, nudust, kappad, adust, gdust, pldust, &
                                        kappav, lTint, nuint, kapBint, kappap) &
                                        bind(C, name='hc_set_dust')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: nlambda(*)
      real(c_double), intent(in) :: nudust(*), kappad(*), adust(*), gdust(*), pldust(*), &
                                    kappav(*), lTint(*), nuint(*), kapBint(*), kappap(*)
    end function hc_set_dust

    integer(c_int) function hc_set_sources(ctx, src) bind(C, name='hc_set_sources')
      import :: c_int, c_ptr, hc_sources
      type(c_ptr), value :: ctx
      type(hc_sources), intent(in) :: src
    end function hc_set_sources

    integer(c_int) function hc_run_iteration(ctx, np, ipeel, seed1, seed2, gen, seg0, out, &
                                             trace) bind(C, name='hc_run_iteration')
      import :: c_int, c_ptr, c_int64_t, c_int32_t, hc_tallies
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: np, seg0
      integer(c_int32_t), value :: ipeel, seed1, seed2, gen
      type(hc_tallies), intent(inout) :: out
      type(c_ptr), value :: trace
    end function hc_run_iteration

    integer(c_int) function hc_temperature(ctx, massarr, ltot, nphot, tdust2, nchanged) &
                                           bind(C, name='hc_temperature')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: massarr(*)
      real(c_double), value :: ltot, nphot
      real(c_double), intent(out) :: tdust2(*)
      integer(c_int64_t), intent(out) :: nchanged
    end function hc_temperature

    ! niter Lucy iterations replayed from one captured HIP graph (the body of
    ! newtts.F90:304-330 without the convergence test)
    integer(c_int) function hc_run_lucy(ctx, niter, np, seed1, seed2, gen, seg0, massarr, &
                                        ltot, out, nchanged) bind(C, name='hc_run_lucy')
      import :: c_int, c_ptr, c_double, c_int64_t, c_int32_t, hc_tallies
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: niter, seed1, seed2, gen
      integer(c_int64_t), value :: np, seg0
      real(c_double), intent(in) :: massarr(*)
      real(c_double), value :: ltot
      type(hc_tallies), intent(inout) :: out
      integer(c_int64_t), intent(out) :: nchanged
    end function hc_run_lucy

    ! niter Lucy iterations of nmodels contexts with all passes in one launch
    ! per kernel variant (a parameter grid); ctxs, np, seg0, ltot, out and
    ! nchanged have nmodels entries; massarr holds nmodels addresses (c_loc of
    ! each model's massarr, or c_null_ptr once a context has it)
    integer(c_int) function hc_run_batch(nmodels, ctxs, niter, np, seed1, seed2, gen, seg0, &
                                         massarr, ltot, oversubscribe, out, nchanged, &
                                         device_ms) bind(C, name='hc_run_batch')
      import :: c_int, c_ptr, c_double, c_int64_t, c_int32_t, hc_tallies
      integer(c_int32_t), value :: nmodels, niter, seed1, seed2, gen, oversubscribe
      type(c_ptr), intent(in) :: ctxs(*)
      integer(c_int64_t), intent(in) :: np(*), seg0(*)
      type(c_ptr), intent(in) :: massarr(*)
      real(c_double), intent(in) :: ltot(*)
      type(hc_tallies), intent(inout) :: out(*)
      integer(c_int64_t), intent(out) :: nchanged(*)
      real(c_double), intent(out) :: device_ms
    end function hc_run_batch

    ! filter curves of filt_mod (after filt) for the filter images
    integer(c_int) function hc_set_filters(ctx, nfilt, nwfiltmax, nwfilt, filtwave, filtphi) &
        bind(C, name='hc_set_filters')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nfilt, nwfiltmax
      integer(c_int32_t), intent(in) :: nwfilt(*)
      real(c_double), intent(in) :: filtwave(*), filtphi(*)
    end function hc_set_filters

    ! image_b_{i,q,u,v} (npeel,nxhst,nxhst,nfilt) of the last peeled pass
    integer(c_int) function hc_get_filter_images(ctx, bi, bq, bu, bv) &
        bind(C, name='hc_get_filter_images')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: bi(*), bq(*), bu(*), bv(*)
    end function hc_get_filter_images

    ! the diffusion approximation (DIFFUSION=YES): gridset's diffus as 0/1 and
    ! diffdir, int8 (nrg,ntg,npg), rarr(nrg), setemiss's kappar(nT,ndg)
    integer(c_int) function hc_set_diffusion(ctx, diffus, diffdir, rarr, kappar) &
        bind(C, name='hc_set_diffusion')
      import :: c_int, c_ptr, c_int8_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int8_t), intent(in) :: diffus(*), diffdir(*)
      real(c_double), intent(in) :: rarr(*), kappar(*)
    end function hc_set_diffusion

    ! Bjorkman-Wood re-emission (CLUCY=NO): kapdBint (nT,nnu,ndg) from setemiss
    ! and massarr (nrg,ntg,npg,ndg); c_null_ptr for both returns to Lucy's
    integer(c_int) function hc_set_bw(ctx, kapdBint, massarr) bind(C, name='hc_set_bw')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, kapdBint, massarr
    end function hc_set_bw

    integer(c_int) function hc_set_option(ctx, name, value) bind(C, name='hc_set_option')
      import :: c_int, c_ptr, c_char, c_int64_t
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)   ! NUL-terminated
      integer(c_int64_t), value :: value
    end function hc_set_option

    ! the stellar spectrum of atmosinit (starbarb.f90:3-118): the atmos_interp
    ! module arrays, e.g. hc_set_atmosphere(hc, nhnu, atmosnu, hnuint, lognu, loghnu)
    integer(c_int) function hc_set_atmosphere(ctx, nhnu, atmosnu, hnuint, lognu, loghnu) &
        bind(C, name='hc_set_atmosphere')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nhnu
      real(c_double), intent(in) :: atmosnu(*), hnuint(*), lognu(*), loghnu(*)
    end function hc_set_atmosphere

    ! the limb-darkening table of `table` (samptabl.f90:1-22): xmu(ntab), prob(ntab)
    integer(c_int) function hc_set_limb_table(ctx, ntab, xmu, prob) &
        bind(C, name='hc_set_limb_table')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ntab
      real(c_double), intent(in) :: xmu(*), prob(*)
    end function hc_set_limb_table

    integer(c_int) function hc_dropped_packets(ctx, pids, cap, ndropped) &
        bind(C, name='hc_dropped_packets')
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: pids(*)
      integer(c_int32_t), value :: cap
      integer(c_int64_t), intent(out) :: ndropped
    end function hc_dropped_packets

    integer(c_int) function hc_destroy(ctx) bind(C, name='hc_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function hc_destroy

    type(c_ptr) function hc_last_error(ctx) bind(C, name='hc_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function hc_last_error

    integer(c_int) function hc_abi_version() bind(C, name='hc_abi_version')
      import :: c_int
    end function hc_abi_version

    ! grid setup on the GPU (gridset.f90:1002-1083): densities, (rho V)^(1/4)
    ! and cell volumes in grid_mod's layout
    integer(c_int) function hc_build_density(model, device, densarr, massarr, vcell) &
                                             bind(C, name='hc_build_density')
      import :: c_int, c_double, hc_density_model
      type(hc_density_model), intent(in) :: model
      integer(c_int), value :: device
      real(c_double), intent(out) :: densarr(*), massarr(*), vcell(*)
    end function hc_build_density

    ! setemiss for one species (dust.f90:394-474): kapBint(nT,nnu), kappap(nT)
    integer(c_int) function hc_build_emissivity(device, nl, nud, kd, ad, nT, lTint, nnu, &
                                                nuint, nq, xq, wq, kapBint, kappap) &
                                                bind(C, name='hc_build_emissivity')
      import :: c_int, c_int32_t, c_double
      integer(c_int), value :: device
      integer(c_int32_t), value :: nl, nT, nnu, nq
      real(c_double), intent(in) :: nud(*), kd(*), ad(*), lTint(*), nuint(*), xq(*), wq(*)
      real(c_double), intent(out) :: kapBint(*), kappap(*)
    end function hc_build_emissivity

    ! clump points per cell of the clumpy-envelope multiplier; pcg holds the
    ! bit patterns of the PCG64 state and increment (lo, hi, lo, hi) and
    ! counts the bit patterns of unsigned 32-bit counts
    integer(c_int) function hc_fractal_counts(device, nrg, ntg, npg, rarr, mcos_theta, &
                                              cos_phi, sin_phi, n, levels, scale, scale_r, &
                                              pcg, counts) bind(C, name='hc_fractal_counts')
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int), value :: device
      integer(c_int32_t), value :: nrg, ntg, npg, n, levels
      real(c_double), intent(in) :: rarr(*), mcos_theta(*), cos_phi(*), sin_phi(*), scale(*)
      real(c_double), value :: scale_r
      integer(c_int64_t), intent(in) :: pcg(4)
      integer(c_int32_t), intent(out) :: counts(*)
    end function hc_fractal_counts
  end interface
end module hc_iface
