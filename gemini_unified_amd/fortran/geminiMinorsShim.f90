! Drop-in replacements of the reference's toroidalMinors, spheroidalMinors and
! unitJumpGreenFunctions modules (f90/toroidalMinors.f90, f90/spheroidalMinors.f90,
! f90/unitJumpGreenFunctions.f90): the same module,
! type and procedure names, the same arguments and the same components filled,
! so that the degree loops that call computeToroidalMinors,
! computeSpheroidalMinors and computeUnitJumpGreenFunctions per (omega, l)
! (computeGreenFKSpectraForSynthetics.f90:419-430, computeGreenFKSpectraForASKI.f90:
! 479-511) compile and run unchanged.
! The minors of one (omega, l) are served from a batch that libgemini_fk.so
! computes on the GPU for a block of consecutive degrees (or wavenumbers) of
! the same frequency, source node range and accuracy: the first call of a
! block fills the cache with gm_minors_batch, later calls copy their degree.
!
! The modules use the reference's geminiIntegrationEnvironment (components
! omre, omim, dll1 of gemini_integration_environment) and errorMessage (add)
! modules, which the host program provides.  gpuMinorsInit hands them the
! context on which gm_model_set and gm_nodes_set were called.
module geminiMinorsCache
   use iso_c_binding
   use geminiFKc
   implicit none
   private
   public :: gpuMinorsInit, gpuMinorsFetch, gpuMinorsNnod, gpuMinorsError, gpuMinorsSolid, gpuUjgfFetch
   public :: c_ut, c_dt, c_dett, c_us, c_ds, c_ulg, c_dlg, c_dets, c_status, c_ujs, c_ujt, c_ustatus
   ! pi of the reference's omega = 2 pi f (mathConstants.f90:28, single precision)
   real(c_double), parameter :: pi32 = real(3.141592653589793, c_double)
   type(c_ptr), save :: ctx = c_null_ptr
   integer, save :: nnod = 0, attmode = 3, block = 512
   integer, allocatable, save :: nstate(:)          ! state of the external nodes: 0 solid, 1 fluid
   real(c_double), save :: rearth = 0.0_c_double, dwn = 0.0_c_double
   ! key of the cached block: omega, source node range, accuracy, first degree and count
   real(c_double), save :: k_omre = -1.0_c_double, k_omim = 0.0_c_double, k_eps = -1.0_c_double
   integer, save :: k_jsl = 0, k_jsr = 0, k_first = 0, k_count = 0
   ! the block's minors, C layout [nnod][comp][k] seen from Fortran as (k, comp, nnod)
   complex(c_double_complex), allocatable, target, save :: c_ut(:,:,:), c_dt(:,:,:), c_dett(:,:)
   complex(c_double_complex), allocatable, target, save :: c_us(:,:,:), c_ds(:,:,:), c_dets(:,:)
   real(c_double), allocatable, target, save :: c_ulg(:,:), c_dlg(:,:)
   integer(c_int), allocatable, target, save :: c_status(:)
   ! the unit-jump Green functions of a block: C [nnod][jump][comp][k] seen as (k, comp, jump, nnod)
   real(c_double), save :: u_omre = -1.0_c_double, u_omim = 0.0_c_double, u_eps = -1.0_c_double
   integer, save :: u_jn = 0, u_first = 0, u_count = 0
   complex(c_double_complex), allocatable, target, save :: c_ujs(:,:,:,:), c_ujt(:,:,:,:)
   integer(c_int), allocatable, target, save :: c_ustatus(:)
   character(len=256), save :: last = ''
contains
   ! the context (model and nodes set), the external nodes' states, earth radius,
   ! attenuation mode, dwn (0: harmonic degrees l with dll1 = l(l+1), > 0:
   ! wavenumbers k = (iwn-1) dwn with dll1 = (k rearth)^2) and the number of
   ! degrees computed per batch
   subroutine gpuMinorsInit(the_ctx, the_state, the_rearth, the_attmode, the_dwn, the_block)
      type(c_ptr), intent(in) :: the_ctx
      integer, intent(in) :: the_state(:), the_attmode
      real(c_double), intent(in) :: the_rearth, the_dwn
      integer, intent(in), optional :: the_block
      ctx = the_ctx; nnod = size(the_state); rearth = the_rearth; attmode = the_attmode; dwn = the_dwn
      if (allocated(nstate)) deallocate(nstate)
      allocate(nstate(nnod))
      nstate = the_state
      if (present(the_block)) block = max(1, the_block)
      k_count = 0; u_count = 0
   end subroutine gpuMinorsInit

   integer function gpuMinorsNnod() result(n)
      n = nnod
   end function gpuMinorsNnod

   logical function gpuMinorsSolid(n) result(solid)
      integer, intent(in) :: n
      solid = nstate(n) == 0
   end function gpuMinorsSolid

   function gpuMinorsError() result(msg)
      character(len=256) :: msg
      msg = last
   end function gpuMinorsError

   ! index k of (omre, omim, dll1) in the cache, after filling it if needed; 0 on an error
   integer function gpuMinorsFetch(omre, omim, dll1, jsl, jsr, eps) result(k)
      real(c_double), intent(in) :: omre, omim, dll1, eps
      integer, intent(in) :: jsl, jsr
      integer :: deg, rc
      integer(c_int), target :: lmin(1), nl(1)
      real(c_double), target :: freq(1)
      integer(c_long_long), target :: stats(GM_NSTAT)
      k = 0
      if (.not. c_associated(ctx)) then
         last = 'gpuMinorsInit was not called'; return
      endif
      deg = degreeOf(dll1)
      if (k_count > 0 .and. omre == k_omre .and. omim == k_omim .and. eps == k_eps .and. jsl == k_jsl .and. &
          jsr == k_jsr .and. deg >= k_first .and. deg < k_first + k_count) then
         k = deg - k_first + 1
         return
      endif
      k_count = 0
      if (allocated(c_ut)) deallocate(c_ut, c_dt, c_dett, c_us, c_ds, c_dets, c_ulg, c_dlg, c_status)
      allocate(c_ut(block, 2, nnod), c_dt(block, 2, nnod), c_dett(block, nnod))
      allocate(c_us(block, 6, nnod), c_ds(block, 6, nnod), c_dets(block, nnod))
      allocate(c_ulg(block, nnod), c_dlg(block, nnod), c_status(block))
      freq(1) = omre / (2.0_c_double * pi32)
      lmin(1) = deg; nl(1) = block
      stats = 0
      if (dwn > 0.0_c_double) then
         rc = gm_minors_batch_wn(ctx, 1_c_int, c_loc(freq), c_loc(lmin), c_loc(nl), dwn, -omim, eps, &
                                 int(attmode, c_int), int(jsl, c_int), int(jsr, c_int), 0_c_int, c_loc(c_ut), &
                                 c_loc(c_dt), c_loc(c_dett), c_loc(c_us), c_loc(c_ds), c_loc(c_ulg), c_loc(c_dlg), &
                                 c_loc(c_dets), c_loc(c_status), c_loc(stats))
      else
         rc = gm_minors_batch(ctx, 1_c_int, c_loc(freq), c_loc(lmin), c_loc(nl), -omim, eps, int(attmode, c_int), &
                              int(jsl, c_int), int(jsr, c_int), 0_c_int, c_loc(c_ut), c_loc(c_dt), c_loc(c_dett), &
                              c_loc(c_us), c_loc(c_ds), c_loc(c_ulg), c_loc(c_dlg), c_loc(c_dets), &
                              c_loc(c_status), c_loc(stats))
      endif
      if (rc /= GM_OK .and. all(c_status == 0)) then
         write(last, '(a,i0)') 'gm_minors_batch failed with status ', rc
         return
      endif
      k_omre = omre; k_omim = omim; k_eps = eps; k_jsl = jsl; k_jsr = jsr; k_first = deg; k_count = block
      k = 1
   end function gpuMinorsFetch

   ! degree (or wavenumber index) of dll1
   integer function degreeOf(dll1) result(deg)
      real(c_double), intent(in) :: dll1
      if (dwn > 0.0_c_double) then
         deg = nint(sqrt(dll1) / (dwn * rearth)) + 1
      else
         deg = nint((-1.0_c_double + sqrt(1.0_c_double + 4.0_c_double * dll1)) / 2.0_c_double)
      endif
   end function degreeOf

   ! index k of (omre, omim, dll1) in the unit-jump Green function cache of a jump
   ! at node jn, after filling it with gm_ujgf_batch if needed; 0 on an error
   integer function gpuUjgfFetch(omre, omim, dll1, jn, eps) result(k)
      real(c_double), intent(in) :: omre, omim, dll1, eps
      integer, intent(in) :: jn
      integer :: deg, rc
      integer(c_int), target :: lmin(1), nl(1)
      real(c_double), target :: freq(1)
      integer(c_long_long), target :: stats(GM_NSTAT)
      k = 0
      if (.not. c_associated(ctx)) then
         last = 'gpuMinorsInit was not called'; return
      endif
      deg = degreeOf(dll1)
      if (u_count > 0 .and. omre == u_omre .and. omim == u_omim .and. eps == u_eps .and. jn == u_jn .and. &
          deg >= u_first .and. deg < u_first + u_count) then
         k = deg - u_first + 1
         return
      endif
      u_count = 0
      if (allocated(c_ujs)) deallocate(c_ujs, c_ujt, c_ustatus)
      allocate(c_ujs(block, 4, 4, nnod), c_ujt(block, 2, 2, nnod), c_ustatus(block))
      freq(1) = omre / (2.0_c_double * pi32)
      lmin(1) = deg; nl(1) = block
      stats = 0
      if (dwn > 0.0_c_double) then
         rc = gm_ujgf_batch_wn(ctx, 1_c_int, c_loc(freq), c_loc(lmin), c_loc(nl), dwn, -omim, eps, &
                               int(attmode, c_int), int(jn, c_int), 0_c_int, c_loc(c_ujs), c_loc(c_ujt), &
                               c_loc(c_ustatus), c_loc(stats))
      else
         rc = gm_ujgf_batch(ctx, 1_c_int, c_loc(freq), c_loc(lmin), c_loc(nl), -omim, eps, int(attmode, c_int), &
                            int(jn, c_int), 0_c_int, c_loc(c_ujs), c_loc(c_ujt), c_loc(c_ustatus), c_loc(stats))
      endif
      if (rc /= GM_OK .and. all(c_ustatus == 0)) then
         write(last, '(a,i0)') 'gm_ujgf_batch failed with status ', rc
         return
      endif
      u_omre = omre; u_omim = omim; u_eps = eps; u_jn = jn; u_first = deg; u_count = block
      k = 1
   end function gpuUjgfFetch
end module geminiMinorsCache

module toroidalMinors
   use geminiIntegrationEnvironment
   use errorMessage
   use geminiMinorsCache
   implicit none
   interface dealloc; module procedure deallocToroidalMinors; end interface
   interface operator (.done.); module procedure doneToroidalMinors; end interface
   type toroidal_minors
      type (gemini_integration_environment) :: intenv
      double complex, dimension(:,:), allocatable :: u_tor             ! upward solutions at nodes (2,jnod)
      double complex, dimension(:,:), allocatable :: d_tor             ! downward solutions at nodes (2,jnod)
      double complex, dimension(:), allocatable :: det                 ! determinant at source nodes
      logical :: done = .false.
   end type
contains
   ! computeToroidalMinors (toroidalMinors.f90:71-258) from the GPU cache
   subroutine computeToroidalMinors(this, jsl, jsr, gem_intenv, eps, secrecy, errmsg)
      type (toroidal_minors) :: this
      integer :: jsl, jsr
      type (gemini_integration_environment) :: gem_intenv
      double precision :: eps
      integer :: secrecy
      type (error_message) :: errmsg
      character(len=21) :: myname = 'computeToroidalMinors'
      integer :: k, n, nn
      nn = gpuMinorsNnod()
      if (nn == 0) then
         call add(errmsg, 2, 'gpuMinorsInit was not called', myname); return
      endif
      if (jsl < 1 .or. jsl > nn .or. jsr < 1 .or. jsr > nn) then
         call add(errmsg, 2, 'Invalid index for source node', myname); return
      endif
      k = gpuMinorsFetch(gem_intenv%omre, gem_intenv%omim, gem_intenv%dll1, jsl, jsr, eps)
      if (k == 0) then
         call add(errmsg, 2, trim(gpuMinorsError()), myname); return
      endif
      if (c_status(k) /= 0) then
         call add(errmsg, 2, 'Toroidal minor integration failed on the GPU', myname); return
      endif
      this%intenv = gem_intenv
      if (allocated(this%u_tor)) deallocate(this%u_tor, this%d_tor, this%det)
      allocate(this%u_tor(2, nn), this%d_tor(2, nn), this%det(nn))
      do n = 1, nn
         this%u_tor(:, n) = c_ut(k, :, n)
         this%d_tor(:, n) = c_dt(k, :, n)
         this%det(n) = c_dett(k, n)
      enddo
      this%done = .true.
   end subroutine computeToroidalMinors

   subroutine deallocToroidalMinors(this)
      type (toroidal_minors) :: this
      if (allocated(this%u_tor)) deallocate(this%u_tor)
      if (allocated(this%d_tor)) deallocate(this%d_tor)
      if (allocated(this%det)) deallocate(this%det)
      this%done = .false.
   end subroutine deallocToroidalMinors

   logical function doneToroidalMinors(this)
      type (toroidal_minors), intent(in) :: this
      doneToroidalMinors = this%done
   end function doneToroidalMinors
end module toroidalMinors

module spheroidalMinors
   use geminiIntegrationEnvironment
   use errorMessage
   use geminiMinorsCache
   implicit none
   interface dealloc; module procedure deallocSpheroidalMinors; end interface
   interface operator (.done.); module procedure doneSpheroidalMinors; end interface
   type spheroidal_minors
      type (gemini_integration_environment) :: intenv
      double complex, dimension(:,:), allocatable :: u_minors          ! upward minors at nodes (6,jnod)
      double complex, dimension(:,:), allocatable :: d_minors          ! downward minors at nodes (6,jnod)
      double complex, dimension(:), allocatable :: det                 ! determinants at nodes
      double complex, dimension(:,:,:), allocatable :: u_mtil          ! upward minors' matrix (4x4) at nodes
      double complex, dimension(:,:,:), allocatable :: d_mtil          ! downward minors' matrix (4x4) at nodes
      double complex, dimension(:,:,:), allocatable :: u_meps          ! upward epsilon*minors' matrix (4x4) at nodes
      double complex, dimension(:,:,:), allocatable :: d_meps          ! downward epsilon*minors' matrix (4x4) at nodes
      double precision, dimension(:), allocatable :: u_lgynorm         ! upward log-normalisation factors at nodes
      double precision, dimension(:), allocatable :: d_lgynorm         ! downward log-normalisation factors at nodes
      logical :: done = .false.
   end type spheroidal_minors
contains
   ! computeSpheroidalMinors (spheroidalMinors.f90:72-585) from the GPU cache; the
   ! minors' matrices of solid nodes (l > 0) follow from the minors as in :306-316,
   ! zero at fluid nodes as there
   subroutine computeSpheroidalMinors(this, jsl, jsr, gem_intenv, eps, secrecy, errmsg)
      type (spheroidal_minors) :: this
      integer :: jsl, jsr
      type (gemini_integration_environment) :: gem_intenv
      double precision :: eps
      integer :: secrecy
      type (error_message) :: errmsg
      character(len=23) :: myname = 'computeSpheroidalMinors'
      integer, dimension(4,4) :: idx
      integer :: k, n, nn, i, j
      double precision :: vf
      nn = gpuMinorsNnod()
      if (nn == 0) then
         call add(errmsg, 2, 'gpuMinorsInit was not called', myname); return
      endif
      if (jsl < 1 .or. jsl > nn .or. jsr < 1 .or. jsr > nn) then
         call add(errmsg, 2, 'Invalid index for source node', myname); return
      endif
      k = gpuMinorsFetch(gem_intenv%omre, gem_intenv%omim, gem_intenv%dll1, jsl, jsr, eps)
      if (k == 0) then
         call add(errmsg, 2, trim(gpuMinorsError()), myname); return
      endif
      if (c_status(k) /= 0) then
         call add(errmsg, 2, 'Spheroidal minor integration failed on the GPU', myname); return
      endif
      this%intenv = gem_intenv
      call deallocSpheroidalMinors(this)
      allocate(this%u_minors(6, nn), this%d_minors(6, nn), this%det(nn), this%u_lgynorm(nn), this%d_lgynorm(nn))
      allocate(this%u_mtil(4, 4, nn), this%d_mtil(4, 4, nn), this%u_meps(4, 4, nn), this%d_meps(4, 4, nn))
      this%u_mtil = (0.d0, 0.d0); this%d_mtil = (0.d0, 0.d0); this%u_meps = (0.d0, 0.d0); this%d_meps = (0.d0, 0.d0)
      ! index of the minor in the 4x4 antisymmetric matrix (upper triangle)
      idx = 0
      idx(1,2) = 1; idx(1,3) = 2; idx(1,4) = 3; idx(2,3) = 4; idx(2,4) = 5; idx(3,4) = 6
      do n = 1, nn
         this%u_minors(:, n) = c_us(k, :, n)
         this%d_minors(:, n) = c_ds(k, :, n)
         this%det(n) = c_dets(k, n)
         this%u_lgynorm(n) = c_ulg(k, n)
         this%d_lgynorm(n) = c_dlg(k, n)
         if (gem_intenv%dll1 > epsilon(1.d0) .and. gpuMinorsSolid(n)) then
            do i = 1, 4
               do j = i + 1, 4
                  vf = -1.d0 + 2.d0 * mod(i + j, 2)
                  this%u_mtil(i, j, n) = this%u_minors(idx(i, j), n)
                  this%u_mtil(j, i, n) = -this%u_mtil(i, j, n)
                  this%u_meps(i, j, n) = vf * this%u_minors(7 - idx(i, j), n)
                  this%u_meps(j, i, n) = -this%u_meps(i, j, n)
                  this%d_mtil(i, j, n) = this%d_minors(idx(i, j), n)
                  this%d_mtil(j, i, n) = -this%d_mtil(i, j, n)
                  this%d_meps(i, j, n) = vf * this%d_minors(7 - idx(i, j), n)
                  this%d_meps(j, i, n) = -this%d_meps(i, j, n)
               enddo
            enddo
         endif
      enddo
      this%done = .true.
   end subroutine computeSpheroidalMinors

   subroutine deallocSpheroidalMinors(this)
      type (spheroidal_minors) :: this
      if (allocated(this%u_minors)) deallocate(this%u_minors)
      if (allocated(this%d_minors)) deallocate(this%d_minors)
      if (allocated(this%det)) deallocate(this%det)
      if (allocated(this%u_mtil)) deallocate(this%u_mtil)
      if (allocated(this%d_mtil)) deallocate(this%d_mtil)
      if (allocated(this%u_meps)) deallocate(this%u_meps)
      if (allocated(this%d_meps)) deallocate(this%d_meps)
      if (allocated(this%u_lgynorm)) deallocate(this%u_lgynorm)
      if (allocated(this%d_lgynorm)) deallocate(this%d_lgynorm)
      this%done = .false.
   end subroutine deallocSpheroidalMinors

   logical function doneSpheroidalMinors(this)
      type (spheroidal_minors), intent(in) :: this
      doneSpheroidalMinors = this%done
   end function doneSpheroidalMinors
end module spheroidalMinors

module unitJumpGreenFunctions
   use geminiIntegrationEnvironment
   use errorMessage
   use toroidalMinors
   use spheroidalMinors
   use geminiMinorsCache
   implicit none
   interface dealloc; module procedure deallocUnitJumpGreenFunctions; end interface
   type unit_jump_green_functions
      double complex, dimension(:,:,:), allocatable :: ujgfsph         ! (4, njs, nnod)
      double complex, dimension(:,:,:), allocatable :: ujgftor         ! (2, njt, nnod), none for a jump in the ocean
   end type
contains
   ! computeUnitJumpGreenFunctions (unitJumpGreenFunctions.f90:62-380) from the GPU
   ! cache: the batch integrates the minors and the spheroidal adjoints of its own,
   ! toromin and spheromin only have to be done, as the reference requires
   subroutine computeUnitJumpGreenFunctions(this, toromin, spheromin, jnods, gem_intenv, eps, secrecy, errmsg)
      type (unit_jump_green_functions) :: this
      type (toroidal_minors) :: toromin
      type (spheroidal_minors) :: spheromin
      integer :: jnods
      type (gemini_integration_environment) :: gem_intenv
      double precision :: eps
      integer :: secrecy
      type (error_message) :: errmsg
      character(len=30) :: myname = 'computeUnitJumpGreenFunctions'
      integer :: k, n, nn, j, njs, njt
      if (.not. .done.toromin) then
         call add(errmsg, 2, 'Toroidal minors not yet computed', myname); return
      endif
      if (.not. .done.spheromin) then
         call add(errmsg, 2, 'Spheroidal minors not yet computed', myname); return
      endif
      nn = gpuMinorsNnod()
      if (jnods < 1 .or. jnods > nn) then
         call add(errmsg, 2, 'Invalid index for source node', myname); return
      endif
      k = gpuUjgfFetch(gem_intenv%omre, gem_intenv%omim, gem_intenv%dll1, jnods, eps)
      if (k == 0) then
         call add(errmsg, 2, trim(gpuMinorsError()), myname); return
      endif
      if (c_ustatus(k) /= 0) then
         call add(errmsg, 2, 'Unit jump Green functions failed on the GPU', myname); return
      endif
      ! a jump in the ocean: (1,0) and (0,1) in (U, R), no toroidal jumps (:118-123)
      njs = 4; njt = 2
      if (.not. gpuMinorsSolid(jnods)) then
         njs = 2; njt = 0
      endif
      call deallocUnitJumpGreenFunctions(this)
      allocate(this%ujgfsph(4, njs, nn))
      if (njt > 0) allocate(this%ujgftor(2, njt, nn))
      do n = 1, nn
         do j = 1, njs
            this%ujgfsph(:, j, n) = c_ujs(k, :, j, n)
         enddo
         do j = 1, njt
            this%ujgftor(:, j, n) = c_ujt(k, :, j, n)
         enddo
      enddo
   end subroutine computeUnitJumpGreenFunctions

   subroutine deallocUnitJumpGreenFunctions(this)
      type (unit_jump_green_functions) :: this
      if (allocated(this%ujgfsph)) deallocate(this%ujgfsph)
      if (allocated(this%ujgftor)) deallocate(this%ujgftor)
   end subroutine deallocUnitJumpGreenFunctions
end module unitJumpGreenFunctions
