! Calls libgemini_fk.so from Fortran through the iso_c_binding modules
! (gemini_unified_amd/fortran), as the reference's drivers would.
!   gfk_call create            create and destroy a context on device 0, print the status
!   gfk_call gfk <in> <out>    model, nodes and one frequency's degree range from <in> (list-directed),
!                              the Green-FK columns (exact variant) into <out> (stream):
!                              rc, nwn, nsp, nnod, gfre(nwn,nsp,nnod), gfim(nwn,nsp,nnod), status(nwn)
program gfk_call
   use iso_c_binding
   use geminiFKc
   use geminiFK
   implicit none
   type(c_ptr) :: ctx
   integer(c_int) :: rc, nk, nlay, nic, noc, fsflag, equidis, nnod, nwn, attmode, jre, istyp, nsp, i
   real(c_double) :: rearth, ric, roc, fref, f, sigma, eps
   real(c_double), allocatable :: rk(:), ro(:), vpv(:), vph(:), vsv(:), vsh(:), eta(:), qkinv(:), qminv(:), rb(:)
   real(c_double), allocatable :: rnod(:)
   integer(c_int), allocatable :: iktop(:), iflso(:), state(:), layer(:), status(:)
   real(c_float), allocatable :: gfre(:,:,:), gfim(:,:,:)
   character(len=512) :: mode, fin, fout
   call get_command_argument(1, mode)
   rc = gm_ctx_create(0_c_int, ctx)
   if (trim(mode) == 'create') then
      print '(a,i0)', 'gm_ctx_create status ', rc
      if (rc /= GM_OK) print '(2a)', 'message: ', gfk_error_string(c_null_ptr)
      if (rc == GM_OK) rc = gm_ctx_destroy(ctx)
      stop
   endif
   if (rc /= GM_OK) then
      print '(2a)', 'no context: ', gfk_error_string(c_null_ptr)
      error stop 2
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
   rc = gm_set_option(ctx, GM_OPT_VARIANT, 0_c_int)
   rc = gfk_model_set(ctx, rk, ro, vpv, vph, vsv, vsh, eta, qkinv, qminv, iktop, rb, iflso, nic, noc, fsflag, &
                      equidis, rearth, ric, roc, fref)
   if (rc /= GM_OK) then
      print '(2a)', 'model: ', gfk_error_string(ctx)
      error stop 3
   endif
   rc = gfk_nodes_set(ctx, rnod, state, layer)
   nsp = 5
   if (istyp == 1) nsp = 10
   allocate(gfre(nwn, nsp, nnod), gfim(nwn, nsp, nnod), status(nwn))
   gfre = 0.0; gfim = 0.0; status = -1
   rc = gfk_frequency(ctx, f, nwn, sigma, eps, attmode, jre, istyp, gfre, gfim, status)
   if (rc /= GM_OK) print '(2a)', 'batch: ', gfk_error_string(ctx)
   open(11, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
   write(11) rc, nwn, nsp, nnod, gfre, gfim, status
   close(11)
   rc = gm_ctx_destroy(ctx)
end program gfk_call
