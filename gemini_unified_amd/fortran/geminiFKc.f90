! iso_c_binding interface of libgemini_fk.so (include/gemini_fk.h).
! Written by gen_interfaces.py from the header; do not edit by hand.
module geminiFKc
   use iso_c_binding
   implicit none
   public
   integer(c_int), parameter :: GM_OK = 0
   integer(c_int), parameter :: GM_ERR_MINSTEP = 1
   integer(c_int), parameter :: GM_ERR_MAXSTEP = 2
   integer(c_int), parameter :: GM_ERR_NODEIDX = 3
   integer(c_int), parameter :: GM_ERR_CF = 4
   integer(c_int), parameter :: GM_ERR_RANGE = 5
   integer(c_int), parameter :: GM_ERR_SPLINE = 6
   integer(c_int), parameter :: GM_ERR_STATE = 7
   integer(c_int), parameter :: GM_ERR_UNSUPPORTED = 8
   integer(c_int), parameter :: GM_ERR_ARG = 9
   integer(c_int), parameter :: GM_ERR_HIP = 10
   integer(c_int), parameter :: GM_ERR_NOMEM = 11
   integer(c_int), parameter :: GM_ELASTIC = 0
   integer(c_int), parameter :: GM_ATTENUATION_ONLY = 1
   integer(c_int), parameter :: GM_DISPERSION_ONLY = 2
   integer(c_int), parameter :: GM_ATTENUATION_AND_DISPERSION = 3
   integer(c_int), parameter :: GM_FLAG_DEVICE_PTRS = 1
   integer(c_int), parameter :: GM_FLAG_STRESS_SPECTRA = 2
   integer(c_int), parameter :: GM_FLAG_DERIV_SPECTRA = 4
   integer(c_int), parameter :: GM_FLAG_OCEAN_SPECTRA = 8
   integer(c_int), parameter :: GM_FILTER_BUTTERWORTH_BANDPASS = 1
   integer(c_int), parameter :: GM_OPT_VARIANT = 1
   integer(c_int), parameter :: GM_OPT_PERSISTENT = 2
   integer(c_int), parameter :: GM_OPT_UNIT = 3
   integer(c_int), parameter :: GM_OPT_WAVES = 4
   integer(c_int), parameter :: GM_OPT_PROPOSE_EVERY = 5
   integer(c_int), parameter :: GM_OPT_SPH_SPLIT = 6
   integer(c_int), parameter :: GM_OPT_ORDER_TARGET = 7
   integer(c_int), parameter :: GM_OPT_GRAPH = 8
   integer(c_int), parameter :: GM_OPT_ADJOINT_PAIR = 9
   integer(c_int), parameter :: GM_NSTAT = 10
   interface
      integer(c_int) function gm_ctx_create(device, out) bind(C, name='gm_ctx_create')
         import
         integer(c_int), value :: device
         type(c_ptr) :: out
      end function gm_ctx_create
      integer(c_int) function gm_ctx_destroy(ctx) bind(C, name='gm_ctx_destroy')
         import
         type(c_ptr), value :: ctx
      end function gm_ctx_destroy
      integer(c_int) function gm_set_option(ctx, option, value) bind(C, name='gm_set_option')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: option
         integer(c_int), value :: value
      end function gm_set_option
      integer(c_int) function gm_last_error(ctx, buf, len) bind(C, name='gm_last_error')
         import
         type(c_ptr), value :: ctx
         character(kind=c_char) :: buf(*)
         integer(c_int), value :: len
      end function gm_last_error
      integer(c_int) function gm_model_set(ctx, nk, rk, ro, vpv, vph, vsv, vsh, eta, qkinv, qminv, nlay, &
          iktop, rb, iflso, nic, noc, fsflag, equidis, rearth, ric, roc, fref) bind(C, name='gm_model_set')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nk
         type(c_ptr), value :: rk
         type(c_ptr), value :: ro
         type(c_ptr), value :: vpv
         type(c_ptr), value :: vph
         type(c_ptr), value :: vsv
         type(c_ptr), value :: vsh
         type(c_ptr), value :: eta
         type(c_ptr), value :: qkinv
         type(c_ptr), value :: qminv
         integer(c_int), value :: nlay
         type(c_ptr), value :: iktop
         type(c_ptr), value :: rb
         type(c_ptr), value :: iflso
         integer(c_int), value :: nic
         integer(c_int), value :: noc
         integer(c_int), value :: fsflag
         integer(c_int), value :: equidis
         real(c_double), value :: rearth
         real(c_double), value :: ric
         real(c_double), value :: roc
         real(c_double), value :: fref
      end function gm_model_set
      integer(c_int) function gm_nodes_set(ctx, nnod, rnod, state, layer) bind(C, name='gm_nodes_set')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nnod
         type(c_ptr), value :: rnod
         type(c_ptr), value :: state
         type(c_ptr), value :: layer
      end function gm_nodes_set
      integer(c_int) function gm_node_properties(ctx, f, attmode, ro, elcon, qinv) bind(C, &
          name='gm_node_properties')
         import
         type(c_ptr), value :: ctx
         real(c_double), value :: f
         integer(c_int), value :: attmode
         type(c_ptr), value :: ro
         type(c_ptr), value :: elcon
         type(c_ptr), value :: qinv
      end function gm_node_properties
      integer(c_int) function gm_minors_batch(ctx, nfreq, freqs, lmin, nl, sigma, eps, attmode, jsl, jsr, &
          flags, u_tor, d_tor, dettor, u_sph, d_sph, ulg, dlg, detsph, status, stats) bind(C, &
          name='gm_minors_batch')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nfreq
         type(c_ptr), value :: freqs
         type(c_ptr), value :: lmin
         type(c_ptr), value :: nl
         real(c_double), value :: sigma
         real(c_double), value :: eps
         integer(c_int), value :: attmode
         integer(c_int), value :: jsl
         integer(c_int), value :: jsr
         integer(c_int), value :: flags
         type(c_ptr), value :: u_tor
         type(c_ptr), value :: d_tor
         type(c_ptr), value :: dettor
         type(c_ptr), value :: u_sph
         type(c_ptr), value :: d_sph
         type(c_ptr), value :: ulg
         type(c_ptr), value :: dlg
         type(c_ptr), value :: detsph
         type(c_ptr), value :: status
         type(c_ptr), value :: stats
      end function gm_minors_batch
      integer(c_int) function gm_gfk_batch(ctx, nfreq, freqs, lmin, nl, sigma, eps, attmode, jre, istyp, &
          flags, gf, gfre, gfim, status, stats) bind(C, name='gm_gfk_batch')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nfreq
         type(c_ptr), value :: freqs
         type(c_ptr), value :: lmin
         type(c_ptr), value :: nl
         real(c_double), value :: sigma
         real(c_double), value :: eps
         integer(c_int), value :: attmode
         integer(c_int), value :: jre
         integer(c_int), value :: istyp
         integer(c_int), value :: flags
         type(c_ptr), value :: gf
         type(c_ptr), value :: gfre
         type(c_ptr), value :: gfim
         type(c_ptr), value :: status
         type(c_ptr), value :: stats
      end function gm_gfk_batch
      integer(c_int) function gm_gfk_batch_noalloc(ctx, nfreq, freqs, lmin, nl, sigma, eps, attmode, jre, &
          istyp, gf, gfre, gfim, status, stats) bind(C, name='gm_gfk_batch_noalloc')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nfreq
         type(c_ptr), value :: freqs
         type(c_ptr), value :: lmin
         type(c_ptr), value :: nl
         real(c_double), value :: sigma
         real(c_double), value :: eps
         integer(c_int), value :: attmode
         integer(c_int), value :: jre
         integer(c_int), value :: istyp
         type(c_ptr), value :: gf
         type(c_ptr), value :: gfre
         type(c_ptr), value :: gfim
         type(c_ptr), value :: status
         type(c_ptr), value :: stats
      end function gm_gfk_batch_noalloc
      integer(c_int) function gm_minors_batch_wn(ctx, nfreq, freqs, iwnmin, nwn, dwn, sigma, eps, attmode, &
          jsl, jsr, flags, u_tor, d_tor, dettor, u_sph, d_sph, ulg, dlg, detsph, status, stats) bind(C, &
          name='gm_minors_batch_wn')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nfreq
         type(c_ptr), value :: freqs
         type(c_ptr), value :: iwnmin
         type(c_ptr), value :: nwn
         real(c_double), value :: dwn
         real(c_double), value :: sigma
         real(c_double), value :: eps
         integer(c_int), value :: attmode
         integer(c_int), value :: jsl
         integer(c_int), value :: jsr
         integer(c_int), value :: flags
         type(c_ptr), value :: u_tor
         type(c_ptr), value :: d_tor
         type(c_ptr), value :: dettor
         type(c_ptr), value :: u_sph
         type(c_ptr), value :: d_sph
         type(c_ptr), value :: ulg
         type(c_ptr), value :: dlg
         type(c_ptr), value :: detsph
         type(c_ptr), value :: status
         type(c_ptr), value :: stats
      end function gm_minors_batch_wn
      integer(c_int) function gm_gfk_batch_wn(ctx, nfreq, freqs, iwnmin, nwn, dwn, sigma, eps, attmode, &
          jre, istyp, flags, gf, gfre, gfim, status, stats) bind(C, name='gm_gfk_batch_wn')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nfreq
         type(c_ptr), value :: freqs
         type(c_ptr), value :: iwnmin
         type(c_ptr), value :: nwn
         real(c_double), value :: dwn
         real(c_double), value :: sigma
         real(c_double), value :: eps
         integer(c_int), value :: attmode
         integer(c_int), value :: jre
         integer(c_int), value :: istyp
         integer(c_int), value :: flags
         type(c_ptr), value :: gf
         type(c_ptr), value :: gfre
         type(c_ptr), value :: gfim
         type(c_ptr), value :: status
         type(c_ptr), value :: stats
      end function gm_gfk_batch_wn
      integer(c_int) function gm_ujgf_batch(ctx, nfreq, freqs, lmin, nl, sigma, eps, attmode, jre, flags, &
          ujs, ujt, status, stats) bind(C, name='gm_ujgf_batch')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nfreq
         type(c_ptr), value :: freqs
         type(c_ptr), value :: lmin
         type(c_ptr), value :: nl
         real(c_double), value :: sigma
         real(c_double), value :: eps
         integer(c_int), value :: attmode
         integer(c_int), value :: jre
         integer(c_int), value :: flags
         type(c_ptr), value :: ujs
         type(c_ptr), value :: ujt
         type(c_ptr), value :: status
         type(c_ptr), value :: stats
      end function gm_ujgf_batch
      integer(c_int) function gm_ujgf_batch_wn(ctx, nfreq, freqs, iwnmin, nwn, dwn, sigma, eps, attmode, &
          jre, flags, ujs, ujt, status, stats) bind(C, name='gm_ujgf_batch_wn')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nfreq
         type(c_ptr), value :: freqs
         type(c_ptr), value :: iwnmin
         type(c_ptr), value :: nwn
         real(c_double), value :: dwn
         real(c_double), value :: sigma
         real(c_double), value :: eps
         integer(c_int), value :: attmode
         integer(c_int), value :: jre
         integer(c_int), value :: flags
         type(c_ptr), value :: ujs
         type(c_ptr), value :: ujt
         type(c_ptr), value :: status
         type(c_ptr), value :: stats
      end function gm_ujgf_batch_wn
      integer(c_int) function gm_aski_batch(ctx, nfreq, freqs, lmin, nl, sigma, eps, attmode, nsnod, jsnod, &
          jre, istyp, flags, gf, gfre, gfim, status, stats) bind(C, name='gm_aski_batch')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nfreq
         type(c_ptr), value :: freqs
         type(c_ptr), value :: lmin
         type(c_ptr), value :: nl
         real(c_double), value :: sigma
         real(c_double), value :: eps
         integer(c_int), value :: attmode
         integer(c_int), value :: nsnod
         type(c_ptr), value :: jsnod
         integer(c_int), value :: jre
         integer(c_int), value :: istyp
         integer(c_int), value :: flags
         type(c_ptr), value :: gf
         type(c_ptr), value :: gfre
         type(c_ptr), value :: gfim
         type(c_ptr), value :: status
         type(c_ptr), value :: stats
      end function gm_aski_batch
      integer(c_int) function gm_last_timing(ctx, ms6) bind(C, name='gm_last_timing')
         import
         type(c_ptr), value :: ctx
         type(c_ptr), value :: ms6
      end function gm_last_timing
      integer(c_int) function gm_last_system_stats(ctx, sys, stats) bind(C, name='gm_last_system_stats')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: sys
         type(c_ptr), value :: stats
      end function gm_last_system_stats
      integer(c_int) function gm_mlp_nparams() bind(C, name='gm_mlp_nparams')
         import
      end function gm_mlp_nparams
      integer(c_int) function gm_mlp_forward(device, params, n, X, Y) bind(C, name='gm_mlp_forward')
         import
         integer(c_int), value :: device
         type(c_ptr), value :: params
         integer(c_int), value :: n
         type(c_ptr), value :: X
         type(c_ptr), value :: Y
      end function gm_mlp_forward
      integer(c_int) function gm_mlp_grad(device, params, n, X, T, w0, w1, loss, grad) bind(C, &
          name='gm_mlp_grad')
         import
         integer(c_int), value :: device
         type(c_ptr), value :: params
         integer(c_int), value :: n
         type(c_ptr), value :: X
         type(c_ptr), value :: T
         real(c_float), value :: w0
         real(c_float), value :: w1
         type(c_ptr), value :: loss
         type(c_ptr), value :: grad
      end function gm_mlp_grad
      integer(c_int) function gm_mlp_train(device, params, n, X, T, w0, w1, epochs, batch, lr, seed, &
          losses) bind(C, name='gm_mlp_train')
         import
         integer(c_int), value :: device
         type(c_ptr), value :: params
         integer(c_int), value :: n
         type(c_ptr), value :: X
         type(c_ptr), value :: T
         real(c_float), value :: w0
         real(c_float), value :: w1
         integer(c_int), value :: epochs
         integer(c_int), value :: batch
         real(c_float), value :: lr
         integer(c_int), value :: seed
         type(c_ptr), value :: losses
      end function gm_mlp_train
      integer(c_int) function gm_set_proposer(ctx, params, norm, safety, kcheck, enable) bind(C, &
          name='gm_set_proposer')
         import
         type(c_ptr), value :: ctx
         type(c_ptr), value :: params
         type(c_ptr), value :: norm
         real(c_float), value :: safety
         integer(c_int), value :: kcheck
         integer(c_int), value :: enable
      end function gm_set_proposer
      integer(c_int) function gm_collect_steplog(ctx, nfreq, freqs, lmin, nl, dwn, sigma, eps, attmode, &
          jsl, jsr, maxrec, records, nrec, stats) bind(C, name='gm_collect_steplog')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nfreq
         type(c_ptr), value :: freqs
         type(c_ptr), value :: lmin
         type(c_ptr), value :: nl
         real(c_double), value :: dwn
         real(c_double), value :: sigma
         real(c_double), value :: eps
         integer(c_int), value :: attmode
         integer(c_int), value :: jsl
         integer(c_int), value :: jsr
         integer(c_int), value :: maxrec
         type(c_ptr), value :: records
         type(c_ptr), value :: nrec
         type(c_ptr), value :: stats
      end function gm_collect_steplog
      integer(c_int) function gm_legendre_batch(ctx, nrcv, theta, lmax, plm, dpdt, pdst, d2pdt2, &
          dtpdst) bind(C, name='gm_legendre_batch')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nrcv
         type(c_ptr), value :: theta
         integer(c_int), value :: lmax
         type(c_ptr), value :: plm
         type(c_ptr), value :: dpdt
         type(c_ptr), value :: pdst
         type(c_ptr), value :: d2pdt2
         type(c_ptr), value :: dtpdst
      end function gm_legendre_batch
      integer(c_int) function gm_moment_sums(ctx, nf1, nf2, nwn, nsp, nkf, gfkspr, gfkspi, nrcv, theta, &
          flags, hdsum) bind(C, name='gm_moment_sums')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nf1
         integer(c_int), value :: nf2
         type(c_ptr), value :: nwn
         integer(c_int), value :: nsp
         integer(c_long_long), value :: nkf
         type(c_ptr), value :: gfkspr
         type(c_ptr), value :: gfkspi
         integer(c_int), value :: nrcv
         type(c_ptr), value :: theta
         integer(c_int), value :: flags
         type(c_ptr), value :: hdsum
      end function gm_moment_sums
      integer(c_int) function gm_force_sums(ctx, nf1, nf2, nwn, nsp, nkf, gfkspr, gfkspi, nrcv, theta, &
          flags, hdsum) bind(C, name='gm_force_sums')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nf1
         integer(c_int), value :: nf2
         type(c_ptr), value :: nwn
         integer(c_int), value :: nsp
         integer(c_long_long), value :: nkf
         type(c_ptr), value :: gfkspr
         type(c_ptr), value :: gfkspi
         integer(c_int), value :: nrcv
         type(c_ptr), value :: theta
         integer(c_int), value :: flags
         type(c_ptr), value :: hdsum
      end function gm_force_sums
      integer(c_int) function gm_synthetics(ctx, nf1, nf2, df, sigma, nwn, nsp, nkf, gfkspr, gfkspi, nrcv, &
          theta, phi, propdir, tred, evfilter, stafilter, istyp, fomt, velocity, nsamp, dt, flags, ts, &
          zsp) bind(C, name='gm_synthetics')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nf1
         integer(c_int), value :: nf2
         real(c_double), value :: df
         real(c_double), value :: sigma
         type(c_ptr), value :: nwn
         integer(c_int), value :: nsp
         integer(c_long_long), value :: nkf
         type(c_ptr), value :: gfkspr
         type(c_ptr), value :: gfkspi
         integer(c_int), value :: nrcv
         type(c_ptr), value :: theta
         type(c_ptr), value :: phi
         type(c_ptr), value :: propdir
         type(c_ptr), value :: tred
         type(c_ptr), value :: evfilter
         type(c_ptr), value :: stafilter
         integer(c_int), value :: istyp
         type(c_ptr), value :: fomt
         integer(c_int), value :: velocity
         integer(c_int), value :: nsamp
         real(c_double), value :: dt
         integer(c_int), value :: flags
         type(c_ptr), value :: ts
         type(c_ptr), value :: zsp
      end function gm_synthetics
      integer(c_int) function gm_strain_spectra(ctx, nf1, nf2, nwn, nsp, nkf, gfkspr, gfkspi, nrcv, theta, &
          phi, re, istyp, fomt, elcon, flags, hdsum, zsp, strainsp, stresssp) bind(C, &
          name='gm_strain_spectra')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nf1
         integer(c_int), value :: nf2
         type(c_ptr), value :: nwn
         integer(c_int), value :: nsp
         integer(c_long_long), value :: nkf
         type(c_ptr), value :: gfkspr
         type(c_ptr), value :: gfkspi
         integer(c_int), value :: nrcv
         type(c_ptr), value :: theta
         type(c_ptr), value :: phi
         real(c_double), value :: re
         integer(c_int), value :: istyp
         type(c_ptr), value :: fomt
         type(c_ptr), value :: elcon
         integer(c_int), value :: flags
         type(c_ptr), value :: hdsum
         type(c_ptr), value :: zsp
         type(c_ptr), value :: strainsp
         type(c_ptr), value :: stresssp
      end function gm_strain_spectra
      integer(c_int) function gm_bessel_sums(ctx, nf1, nf2, nwn, nsp, nkf, gfkspr, gfkspi, dwn, rearth, &
          nrcv, x, flags, wnint) bind(C, name='gm_bessel_sums')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nf1
         integer(c_int), value :: nf2
         type(c_ptr), value :: nwn
         integer(c_int), value :: nsp
         integer(c_long_long), value :: nkf
         type(c_ptr), value :: gfkspr
         type(c_ptr), value :: gfkspi
         real(c_double), value :: dwn
         real(c_double), value :: rearth
         integer(c_int), value :: nrcv
         type(c_ptr), value :: x
         integer(c_int), value :: flags
         type(c_ptr), value :: wnint
      end function gm_bessel_sums
      integer(c_int) function gm_force_bessel_sums(ctx, nf1, nf2, nwn, nsp, nkf, gfkspr, gfkspi, dwn, &
          rearth, nrcv, x, flags, wnint) bind(C, name='gm_force_bessel_sums')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nf1
         integer(c_int), value :: nf2
         type(c_ptr), value :: nwn
         integer(c_int), value :: nsp
         integer(c_long_long), value :: nkf
         type(c_ptr), value :: gfkspr
         type(c_ptr), value :: gfkspi
         real(c_double), value :: dwn
         real(c_double), value :: rearth
         integer(c_int), value :: nrcv
         type(c_ptr), value :: x
         integer(c_int), value :: flags
         type(c_ptr), value :: wnint
      end function gm_force_bessel_sums
      integer(c_int) function gm_strain_spectra_regional(ctx, nf1, nf2, nwn, nsp, nkf, gfkspr, gfkspi, &
          nrcv, x, phi, re, dwn, rearth, istyp, fomt, elcon, flags, hdsum, zsp, strainsp, stresssp) bind(C, &
          name='gm_strain_spectra_regional')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nf1
         integer(c_int), value :: nf2
         type(c_ptr), value :: nwn
         integer(c_int), value :: nsp
         integer(c_long_long), value :: nkf
         type(c_ptr), value :: gfkspr
         type(c_ptr), value :: gfkspi
         integer(c_int), value :: nrcv
         type(c_ptr), value :: x
         type(c_ptr), value :: phi
         real(c_double), value :: re
         real(c_double), value :: dwn
         real(c_double), value :: rearth
         integer(c_int), value :: istyp
         type(c_ptr), value :: fomt
         type(c_ptr), value :: elcon
         integer(c_int), value :: flags
         type(c_ptr), value :: hdsum
         type(c_ptr), value :: zsp
         type(c_ptr), value :: strainsp
         type(c_ptr), value :: stresssp
      end function gm_strain_spectra_regional
      integer(c_int) function gm_synthetics_regional(ctx, nf1, nf2, df, sigma, nwn, nsp, nkf, gfkspr, &
          gfkspi, dwn, rearth, nrcv, x, phi, propdir, tred, evfilter, stafilter, istyp, fomt, velocity, &
          nsamp, dt, flags, ts, zsp) bind(C, name='gm_synthetics_regional')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nf1
         integer(c_int), value :: nf2
         real(c_double), value :: df
         real(c_double), value :: sigma
         type(c_ptr), value :: nwn
         integer(c_int), value :: nsp
         integer(c_long_long), value :: nkf
         type(c_ptr), value :: gfkspr
         type(c_ptr), value :: gfkspi
         real(c_double), value :: dwn
         real(c_double), value :: rearth
         integer(c_int), value :: nrcv
         type(c_ptr), value :: x
         type(c_ptr), value :: phi
         type(c_ptr), value :: propdir
         type(c_ptr), value :: tred
         type(c_ptr), value :: evfilter
         type(c_ptr), value :: stafilter
         integer(c_int), value :: istyp
         type(c_ptr), value :: fomt
         integer(c_int), value :: velocity
         integer(c_int), value :: nsamp
         real(c_double), value :: dt
         integer(c_int), value :: flags
         type(c_ptr), value :: ts
         type(c_ptr), value :: zsp
      end function gm_synthetics_regional
      integer(c_int) function gm_event_filter(filtype, filspecs, nf1, nf2, df, sigma, hfil) bind(C, &
          name='gm_event_filter')
         import
         integer(c_int), value :: filtype
         type(c_ptr), value :: filspecs
         integer(c_int), value :: nf1
         integer(c_int), value :: nf2
         real(c_double), value :: df
         real(c_double), value :: sigma
         type(c_ptr), value :: hfil
      end function gm_event_filter
      integer(c_int) function gm_travel_times(ctx, nnod, ntp, istele, rnod, rtp, slowness, delta, ttime, &
          rs, nrcv, re, deltain, tt, slow, status) bind(C, name='gm_travel_times')
         import
         type(c_ptr), value :: ctx
         integer(c_int), value :: nnod
         integer(c_int), value :: ntp
         integer(c_int), value :: istele
         type(c_ptr), value :: rnod
         type(c_ptr), value :: rtp
         type(c_ptr), value :: slowness
         type(c_ptr), value :: delta
         type(c_ptr), value :: ttime
         real(c_double), value :: rs
         integer(c_int), value :: nrcv
         type(c_ptr), value :: re
         type(c_ptr), value :: deltain
         type(c_ptr), value :: tt
         type(c_ptr), value :: slow
         type(c_ptr), value :: status
      end function gm_travel_times
      integer(c_int) function gm_last_synth_timing(ctx, ms6) bind(C, name='gm_last_synth_timing')
         import
         type(c_ptr), value :: ctx
         type(c_ptr), value :: ms6
      end function gm_last_synth_timing
   end interface
end module geminiFKc
