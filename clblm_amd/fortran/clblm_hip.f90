! ISO_C_BINDING interface to libclblm_hip.so (include/clblm_hip.h).
!
! clblm_hip_odlay_plain is the per-layer optical-depth call with plain arguments
! (HITRAN molecule numbers and columns, FP64 throughout); Module_ODLAY in
! module_odlay_hip.f90 builds the reference's ODLAY entry points on top of it.
module clblm_hip
   use iso_c_binding
   implicit none
   private
   public :: clblm_layer_t, clblm_od_params_t, CLBLM_FP64, CLBLM_FP32C
   public :: clblm_hip_init, clblm_hip_lines, clblm_hip_odlay_plain, clblm_hip_finalize, clblm_hip_error
   public :: clblm_hip_jrad

   integer(c_int), parameter :: CLBLM_FP64 = 0, CLBLM_FP32C = 1
   integer, parameter :: MAXMOL = 47

   type, bind(C) :: clblm_layer_t
      real(c_double) :: pave, tave, wtot, zbot, ztop
      real(c_double) :: w(0:MAXMOL)
      integer(c_int32_t) :: sel(0:MAXMOL)
   end type

   type, bind(C) :: clblm_od_params_t
      real(c_double) :: v1, v2, dvnormal, dvnarrow, sample, dvout
      integer(c_int32_t) :: precision, contnm
      type(c_ptr) :: continuum
      integer(c_int32_t) :: nptabs, flags
   end type

   interface
      integer(c_int) function c_ctx_create(device, ctx) bind(C, name="clblm_ctx_create")
         import :: c_int, c_ptr
         integer(c_int), value :: device
         type(c_ptr), intent(out) :: ctx
      end function
      integer(c_int) function c_ctx_destroy(ctx) bind(C, name="clblm_ctx_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      type(c_ptr) function c_last_error() bind(C, name="clblm_last_error")
         import :: c_ptr
      end function
      integer(c_int) function c_upload(ctx, n, nu, s, a_air, a_self, e_lo, tmpalf, shift, mol, iflg, handle) &
            bind(C, name="clblm_linelist_upload")
         import :: c_int, c_ptr, c_int64_t, c_double, c_float, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: n
         real(c_double), intent(in) :: nu(*)
         real(c_float), intent(in) :: s(*), a_air(*), a_self(*), e_lo(*), tmpalf(*), shift(*)
         integer(c_int32_t), intent(in) :: mol(*), iflg(*)
         integer(c_int), intent(out) :: handle
      end function
      integer(c_int) function c_odlay(ctx, handle, layer, par, jrad, nsamp, dv_used, od) bind(C, name="clblm_odlay")
         import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double, clblm_layer_t, clblm_od_params_t
         type(c_ptr), value :: ctx
         integer(c_int), value :: handle
         type(clblm_layer_t), intent(in) :: layer
         type(clblm_od_params_t), intent(in) :: par
         integer(c_int32_t), intent(inout) :: jrad
         integer(c_int64_t), intent(out) :: nsamp
         real(c_double), intent(out) :: dv_used
         type(c_ptr), value :: od
      end function
   end interface

   type(c_ptr), save :: g_ctx = c_null_ptr
   integer(c_int), save :: g_handle = -1
   ! the reference's SAVEd JRAD (clblm_ODLAY.f90:182-183): starts at the control value
   integer(c_int32_t), save :: clblm_hip_jrad = 1

contains

   subroutine clblm_hip_error(rc, where)
      integer(c_int), intent(in) :: rc
      character(*), intent(in) :: where
      character(kind=c_char), pointer :: msg(:)
      integer :: n
      if (rc == 0) return
      call c_f_pointer(c_last_error(), msg, [512])
      n = 0
      do while (n < 512)
         if (msg(n + 1) == c_null_char) exit
         n = n + 1
      end do
      write (*, '(3a,i0,2a)') '--- ', where, '(): clblm error ', rc, ': ', transfer(msg(1:n), repeat(' ', n))
      stop 1
   end subroutine

   subroutine clblm_hip_init(device)
      integer, intent(in), optional :: device
      integer(c_int) :: dev
      dev = 0
      if (present(device)) dev = device
      if (c_associated(g_ctx)) return
      call clblm_hip_error(c_ctx_create(dev, g_ctx), 'clblm_hip_init')
   end subroutine

   ! Upload the line list once (replaces the TAPE3 reads of RDLIN and RDLNFL).
   subroutine clblm_hip_lines(nu, s, a_air, a_self, e_lo, tmpalf, shift, mol, iflg)
      real(c_double), intent(in) :: nu(:)
      real(c_float), intent(in) :: s(:), a_air(:), a_self(:), e_lo(:), tmpalf(:), shift(:)
      integer(c_int32_t), intent(in) :: mol(:), iflg(:)
      call clblm_hip_init()
      call clblm_hip_error(c_upload(g_ctx, int(size(nu), c_int64_t), nu, s, a_air, a_self, e_lo, tmpalf, &
                                    shift, mol, iflg, g_handle), 'clblm_hip_lines')
   end subroutine

   ! One layer: molecule numbers mols(1:n) with columns w(1:n), p (hPa), T (K), Wtot.
   subroutine clblm_hip_odlay_plain(mols, w, pave, tave, wtot, v1, v2, dvnormal, dvnarrow, sample, dvout, &
                                    precision, nsamp, od, dv_used, continuum)
      integer, intent(in) :: mols(:)
      real(c_double), intent(in) :: w(:), pave, tave, wtot, v1, v2, dvnormal, dvnarrow, sample, dvout
      integer(c_int), intent(in) :: precision
      integer, intent(out) :: nsamp
      real(c_double), allocatable, target, intent(out) :: od(:)
      real(c_double), intent(out) :: dv_used
      real(c_double), intent(in), optional, target :: continuum(:)
      type(clblm_layer_t) :: lay
      type(clblm_od_params_t) :: par
      integer(c_int32_t) :: jr
      integer(c_int64_t) :: ns
      integer :: k
      if (g_handle < 0) stop '--- clblm_hip_odlay_plain(): no line list uploaded (call clblm_hip_lines)'
      lay%pave = pave; lay%tave = tave; lay%wtot = wtot; lay%zbot = 0; lay%ztop = 0
      lay%w = 0; lay%sel = 0
      do k = 1, size(mols)
         lay%w(mols(k)) = w(k)
         lay%sel(mols(k)) = 1
      end do
      par%v1 = v1; par%v2 = v2; par%dvnormal = dvnormal; par%dvnarrow = dvnarrow; par%sample = sample
      par%dvout = dvout; par%precision = precision; par%contnm = 0; par%continuum = c_null_ptr
      par%nptabs = 0; par%flags = 0
      if (present(continuum)) then
         par%contnm = 1
         par%continuum = c_loc(continuum)
         par%nptabs = size(continuum)
      end if
      jr = clblm_hip_jrad
      call clblm_hip_error(c_odlay(g_ctx, g_handle, lay, par, jr, ns, dv_used, c_null_ptr), 'clblm_hip_odlay')
      allocate (od(ns))
      jr = clblm_hip_jrad
      call clblm_hip_error(c_odlay(g_ctx, g_handle, lay, par, jr, ns, dv_used, c_loc(od)), 'clblm_hip_odlay')
      clblm_hip_jrad = jr
      nsamp = int(ns)
   end subroutine

   subroutine clblm_hip_finalize()
      integer(c_int) :: rc
      if (c_associated(g_ctx)) rc = c_ctx_destroy(g_ctx)
      g_ctx = c_null_ptr
      g_handle = -1
   end subroutine
end module
