! The degree loop of the reference's synthetics driver around the drop-in
! toroidalMinors / spheroidalMinors modules (computeGreenFKSpectraForSynthetics.f90:
! 407-425): per degree, set the integration environment and call
! computeToroidalMinors and computeSpheroidalMinors with the receiver as the
! source node range; the minors come from the GPU cache.
!   minors_shim nocontext         the shims before gpuMinorsInit: a level-2 error, printed
!   minors_shim minors <in> <out> model, nodes and one frequency from <in> (gfk_call's format);
!                                 per degree l = 0..nwn-1 into <out> (stream): level, u_tor(2,nnod),
!                                 d_tor, det_tor(jre), u_minors(6,nnod), d_minors, det_sph(jre),
!                                 u_mtil(4,4,nnod), d_meps(4,4,nnod), ujgfsph(4,4,nnod), ujgftor(2,2,nnod)
program minors_shim
   use iso_c_binding
   use geminiFKc
   use geminiFK
   use geminiIntegrationEnvironment
   use errorMessage
   use geminiMinorsCache
   use toroidalMinors
   use spheroidalMinors
   use unitJumpGreenFunctions
   implicit none
   type(c_ptr) :: ctx
   type (gemini_integration_environment) :: env
   type (error_message) :: errmsg
   type (toroidal_minors) :: toromin
   type (spheroidal_minors) :: spheromin
   type (unit_jump_green_functions) :: ujgf
   integer(c_int) :: rc, nk, nlay, nic, noc, fsflag, equidis, nnod, nwn, attmode, jre, istyp, i, l
   real(c_double) :: rearth, ric, roc, fref, f, sigma, eps, wn
   real(c_double), allocatable :: rk(:), ro(:), vpv(:), vph(:), vsv(:), vsh(:), eta(:), qkinv(:), qminv(:), rb(:)
   real(c_double), allocatable :: rnod(:)
   integer(c_int), allocatable :: iktop(:), iflso(:), state(:), layer(:)
   character(len=512) :: mode, fin, fout
   call get_command_argument(1, mode)
   if (trim(mode) == 'nocontext') then
      env%omre = 0.1d0; env%dll1 = 2.d0
      call computeToroidalMinors(toromin, 1, 1, env, 1.d-6, 0, errmsg)
      print '(a,i0,2a)', 'level ', .level.errmsg, ': ', trim(errmsg%message)
      stop
   endif
   call get_command_argument(2, fin)
   call get_command_argument(3, fout)
   open(10, file=trim(fin), status='old', action='read')
   read(10, *) nk, nlay, nic, noc, fsflag, equidis
   read(10, *) rearth, ric, roc, fref
   allocate(rk(nk), ro(nk), vpv(nk), vph(nk), vsv(nk), vsh(nk), eta(nk), qkinv(nk), qminv(nk))
   do i = 1, nk
      read(10, *) rk(i), ro(i), vpv(i), vph(i), vsv(i), vsh(i), eta(i), qkinv(i), qminv(i)
   enddo
   allocate(iktop(nlay), rb(nlay), iflso(nlay))
   do i = 1, nlay
      read(10, *) iktop(i), rb(i), iflso(i)
   enddo
   read(10, *) nnod
   allocate(rnod(nnod), state(nnod), layer(nnod))
   do i = 1, nnod
      read(10, *) rnod(i), state(i), layer(i)
   enddo
   read(10, *) f, nwn, sigma, eps, attmode, jre, istyp
   close(10)
   rc = gm_ctx_create(0_c_int, ctx)
   if (rc /= GM_OK) then
      print '(2a)', 'no context: ', gfk_error_string(c_null_ptr)
      error stop 2
   endif
   rc = gm_set_option(ctx, GM_OPT_VARIANT, 0_c_int)
   rc = gfk_model_set(ctx, rk, ro, vpv, vph, vsv, vsh, eta, qkinv, qminv, iktop, rb, iflso, nic, noc, fsflag, &
                      equidis, rearth, ric, roc, fref)
   if (rc == GM_OK) rc = gfk_nodes_set(ctx, rnod, state, layer)
   if (rc /= GM_OK) then
      print '(2a)', 'model or nodes: ', gfk_error_string(ctx)
      error stop 3
   endif
   ! 16 degrees per batch: the loop below refills the cache several times
   call gpuMinorsInit(ctx, int(state), rearth, int(attmode), 0.0_c_double, 16)
   open(11, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
   do l = 0, nwn - 1
      ! createGemini's omega and degree (computeGreenFKSpectraForSynthetics.f90:391, 411-418)
      env%omre = 2.d0 * real(3.141592653589793, c_double) * f
      env%omim = -sigma
      wn = sqrt(dble(l * (l + 1))) / rearth
      env%dll1 = (wn * rearth)**2
      call computeToroidalMinors(toromin, jre, jre, env, eps, 0, errmsg)
      if (.level.errmsg /= 2) call computeSpheroidalMinors(spheromin, jre, jre, env, eps, 0, errmsg)
      if (.level.errmsg /= 2) call computeUnitJumpGreenFunctions(ujgf, toromin, spheromin, jre, env, eps, 0, errmsg)
      write(11) .level.errmsg
      if (.level.errmsg == 2) exit
      write(11) toromin%u_tor, toromin%d_tor, toromin%det(jre), spheromin%u_minors, spheromin%d_minors, &
                spheromin%det(jre), spheromin%u_mtil, spheromin%d_meps, ujgf%ujgfsph, ujgf%ujgftor
      call dealloc(toromin)
      call dealloc(spheromin)
      call dealloc(ujgf)
   enddo
   close(11)
   rc = gm_ctx_destroy(ctx)
end program minors_shim
