! Fortran-side helpers over the C ABI (geminiFKc): the last error as a Fortran
! string, the model and node upload from the reference's 1-based arrays, and
! the per-frequency Green-FK batch that replaces the degree loop body of
! computeGreenFKSpectraForSynthetics (f90/computeGreenFKSpectraForSynthetics.f90:404-494):
! gfre/gfim(nwn, nsp, nnod) of one frequency, l = 0 .. nwn-1 (global) or
! wavenumber indices 2 .. nwn (regional, column 1 left zero as there).
module geminiFK
   use iso_c_binding
   use geminiFKc
   implicit none
   private
   public :: gfk_error_string, gfk_model_set, gfk_nodes_set, gfk_frequency, gfk_frequency_wn
contains

   function gfk_error_string(ctx) result(msg)
      type(c_ptr), intent(in) :: ctx
      character(len=:), allocatable :: msg
      character(kind=c_char) :: buf(512)
      integer :: i, n, rc
      buf = c_null_char
      rc = gm_last_error(ctx, buf, 512_c_int)
      n = 0
      do i = 1, 512
         if (buf(i) == c_null_char) exit
         n = i
      enddo
      allocate(character(len=n) :: msg)
      do i = 1, n
         msg(i:i) = buf(i)
      enddo
   end function gfk_error_string

   ! node earth model in SI units, knots bottom-up (createNodeEarthmodel's arrays)
   integer(c_int) function gfk_model_set(ctx, rk, ro, vpv, vph, vsv, vsh, eta, qkinv, qminv, iktop, rb, iflso, &
                                         nic, noc, fsflag, equidis, rearth, ric, roc, fref) result(rc)
      type(c_ptr), intent(in) :: ctx
      real(c_double), intent(in), target, contiguous :: rk(:), ro(:), vpv(:), vph(:), vsv(:), vsh(:), eta(:), qkinv(:), &
                                                      qminv(:)
      real(c_double), intent(in), target, contiguous :: rb(:)
      integer(c_int), intent(in), target, contiguous :: iktop(:), iflso(:)
      integer(c_int), intent(in) :: nic, noc, fsflag, equidis
      real(c_double), intent(in) :: rearth, ric, roc, fref
      rc = gm_model_set(ctx, int(size(rk), c_int), c_loc(rk), c_loc(ro), c_loc(vpv), c_loc(vph), c_loc(vsv), &
                        c_loc(vsh), c_loc(eta), c_loc(qkinv), c_loc(qminv), int(size(rb), c_int), c_loc(iktop), &
                        c_loc(rb), c_loc(iflso), nic, noc, fsflag, equidis, rearth, ric, roc, fref)
   end function gfk_model_set

   ! external radial nodes: ascending radii, state 1 fluid / 0 solid, 1-based layer
   integer(c_int) function gfk_nodes_set(ctx, rnod, state, layer) result(rc)
      type(c_ptr), intent(in) :: ctx
      real(c_double), intent(in), target, contiguous :: rnod(:)
      integer(c_int), intent(in), target, contiguous :: state(:), layer(:)
      rc = gm_nodes_set(ctx, int(size(rnod), c_int), c_loc(rnod), c_loc(state), c_loc(layer))
   end function gfk_nodes_set

   ! global mode: one frequency f, degrees l = 0 .. nwn-1 (the reference's iwn = 1 .. nwn);
   ! gfre, gfim(nwn, nsp, nnod), status(nwn)
   integer(c_int) function gfk_frequency(ctx, f, nwn, sigma, eps, attmode, jre, istyp, gfre, gfim, status) result(rc)
      type(c_ptr), intent(in) :: ctx
      real(c_double), intent(in) :: f, sigma, eps
      integer(c_int), intent(in) :: nwn, attmode, jre, istyp
      real(c_float), intent(inout), target, contiguous :: gfre(:,:,:), gfim(:,:,:)
      integer(c_int), intent(inout), target, contiguous :: status(:)
      real(c_double), target :: freqs(1)
      integer(c_int), target :: lmin(1), nl(1)
      freqs(1) = f; lmin(1) = 0; nl(1) = nwn
      rc = gm_gfk_batch(ctx, 1_c_int, c_loc(freqs), c_loc(lmin), c_loc(nl), sigma, eps, attmode, jre, istyp, &
                        0_c_int, c_null_ptr, c_loc(gfre), c_loc(gfim), c_loc(status), c_null_ptr)
   end function gfk_frequency

   ! regional mode: wavenumber indices iwn = 2 .. nwn, k = (iwn-1) dwn; gfre, gfim(nwn, nsp, nnod) with
   ! column iwn = 1 (k = 0) zero as the reference writes it; status(nwn-1)
   integer(c_int) function gfk_frequency_wn(ctx, f, nwn, dwn, sigma, eps, attmode, jre, istyp, gfre, gfim, status) &
                                            result(rc)
      type(c_ptr), intent(in) :: ctx
      real(c_double), intent(in) :: f, dwn, sigma, eps
      integer(c_int), intent(in) :: nwn, attmode, jre, istyp
      real(c_float), intent(inout), target, contiguous :: gfre(:,:,:), gfim(:,:,:)
      integer(c_int), intent(inout), target, contiguous :: status(:)
      real(c_double), target :: freqs(1)
      integer(c_int), target :: iwn0(1), n(1)
      real(c_float), allocatable, target :: br(:,:,:), bi(:,:,:)
      integer :: isp, jn
      freqs(1) = f; iwn0(1) = 2; n(1) = nwn - 1
      allocate(br(nwn - 1, size(gfre, 2), size(gfre, 3)), bi(nwn - 1, size(gfre, 2), size(gfre, 3)))
      rc = gm_gfk_batch_wn(ctx, 1_c_int, c_loc(freqs), c_loc(iwn0), c_loc(n), dwn, sigma, eps, attmode, jre, istyp, &
                           0_c_int, c_null_ptr, c_loc(br), c_loc(bi), c_loc(status), c_null_ptr)
      do jn = 1, size(gfre, 3)
         do isp = 1, size(gfre, 2)
            gfre(1, isp, jn) = 0.0; gfim(1, isp, jn) = 0.0
            gfre(2:nwn, isp, jn) = br(:, isp, jn); gfim(2:nwn, isp, jn) = bi(:, isp, jn)
         enddo
      enddo
   end function gfk_frequency_wn

end module geminiFK
