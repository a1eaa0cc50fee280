! ISO_C_BINDING interface to libclblm_hip.so (include/clblm_hip.h).
!
! clblm_hip_odlay_plain is the per-layer optical-depth call with plain arguments
! (HITRAN molecule numbers and columns); Module_ODLAY in module_odlay_hip.f90 builds
! the reference's ODLAY entry points on top of it.  The line list comes from a TAPE3
! file (clblm_hip_lines_file) or from arrays (clblm_hip_lines); the MT_CKD water-vapour
! continuum table from the reference's netCDF file (clblm_hip_mtckd).  The batched
! optical depths, the radiance/Jacobian pass and the level Jacobians take device
! arrays, which Fortran callers hold through clblm_hip_dev_*.
module clblm_hip
   use iso_c_binding
   implicit none
   private
   public :: clblm_layer_t, clblm_od_params_t, CLBLM_FP64, CLBLM_FP32C, CLBLM_OD_SPECIES_BROAD, CLBLM_OD_MTCKD
   public :: CLBLM_OD_CONTNM
   public :: clblm_hip_init, clblm_hip_lines, clblm_hip_lines_file, clblm_hip_have_lines, clblm_hip_mtckd
   public :: clblm_hip_contnm_scales
   public :: clblm_hip_odlay_plain, clblm_hip_finalize, clblm_hip_error, clblm_hip_jrad, clblm_hip_precision
   public :: clblm_hip_ctx, clblm_hip_handle
   public :: clblm_hip_dev_alloc, clblm_hip_dev_free, clblm_hip_dev_put, clblm_hip_dev_get
   public :: c_od_batch, c_rt_jac, c_level_jac

   integer(c_int), parameter :: CLBLM_FP64 = 0, CLBLM_FP32C = 1
   integer(c_int32_t), parameter :: CLBLM_OD_SPECIES_BROAD = 1, CLBLM_OD_MTCKD = 2, CLBLM_OD_CONTNM = 4
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
      integer(c_int) function c_load_tape3(ctx, path, vlo, vhi, handle, n) bind(C, name="clblm_linelist_load_tape3")
         import :: c_int, c_ptr, c_char, c_double, c_int64_t
         type(c_ptr), value :: ctx
         character(kind=c_char), intent(in) :: path(*)
         real(c_double), value :: vlo, vhi
         integer(c_int), intent(out) :: handle
         integer(c_int64_t), intent(out) :: n
      end function
      integer(c_int) function c_mtckd_load_nc(ctx, path, xself, xfrgn) bind(C, name="clblm_mtckd_load_nc")
         import :: c_int, c_ptr, c_char, c_double
         type(c_ptr), value :: ctx
         character(kind=c_char), intent(in) :: path(*)
         real(c_double), value :: xself, xfrgn
      end function
      integer(c_int) function c_contnm_scales(ctx, xco2c, xo3cn, xo2cn, xn2cn, xrayl) &
            bind(C, name="clblm_contnm_scales")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: xco2c, xo3cn, xo2cn, xn2cn, xrayl
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
      ! many layers into one device array (d_od == c_null_ptr: the offsets only)
      integer(c_int) function c_od_batch(ctx, handle, nlay, layers, pars, jrad, offsets, d_od, stream) &
            bind(C, name="clblm_od_batch")
         import :: c_int, c_ptr, c_int32_t, c_int64_t, clblm_layer_t, clblm_od_params_t
         type(c_ptr), value :: ctx
         integer(c_int), value :: handle, nlay
         type(clblm_layer_t), intent(in) :: layers(*)
         type(clblm_od_params_t), intent(in) :: pars(*)
         integer(c_int32_t), intent(inout) :: jrad
         integer(c_int64_t), intent(out) :: offsets(*)
         type(c_ptr), value :: d_od, stream
      end function
      ! radiance and its per-point derivatives (device arrays)
      integer(c_int) function c_rt_jac(ctx, nlay, npts, v1, dv, d_od, d_tlay, d_tlev, d_dt_dtup, d_dt_dtlo, &
                                       surf_emis, t_skin, lin_in_tau, sec_dn, d_rad, d_dr_dtau, d_dr_dtlev, &
                                       d_dr_dtskin, d_dr_demis, stream) bind(C, name="clblm_rt_jac")
         import :: c_int, c_ptr, c_int64_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: nlay, lin_in_tau
         integer(c_int64_t), value :: npts
         real(c_double), value :: v1, dv, surf_emis, t_skin, sec_dn
         type(c_ptr), value :: d_od, d_tlay, d_tlev, d_dt_dtup, d_dt_dtlo, d_rad, d_dr_dtau, d_dr_dtlev
         type(c_ptr), value :: d_dr_dtskin, d_dr_demis, stream
      end function
      ! level Jacobians from the layer-OD variants (device arrays)
      integer(c_int) function c_level_jac(ctx, nlay, nmol, npts, d_od_var, d_dr_dtau, d_dr_dtlev, d_wtot, d_tlay, &
                                          dt, dw_frac, d_dt_dtup, d_dt_dtlo, d_dw_dqup, d_dw_dqlo, d_jac_t, &
                                          d_jac_q, stream) bind(C, name="clblm_level_jac")
         import :: c_int, c_ptr, c_int64_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: nlay, nmol
         integer(c_int64_t), value :: npts
         real(c_double), value :: dt, dw_frac
         type(c_ptr), value :: d_od_var, d_dr_dtau, d_dr_dtlev, d_wtot, d_tlay, d_dt_dtup, d_dt_dtlo
         type(c_ptr), value :: d_dw_dqup, d_dw_dqlo, d_jac_t, d_jac_q, stream
      end function
      integer(c_int) function c_dev_alloc(ctx, bytes, d) bind(C, name="clblm_dev_alloc")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: bytes
         type(c_ptr), intent(out) :: d
      end function
      integer(c_int) function c_dev_free(ctx, d) bind(C, name="clblm_dev_free")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, d
      end function
      integer(c_int) function c_dev_put(ctx, d, h, bytes) bind(C, name="clblm_dev_put")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx, d, h
         integer(c_int64_t), value :: bytes
      end function
      integer(c_int) function c_dev_get(ctx, h, d, bytes) bind(C, name="clblm_dev_get")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx, h, d
         integer(c_int64_t), value :: bytes
      end function
   end interface

   type(c_ptr), save :: clblm_hip_ctx = c_null_ptr
   integer(c_int), save :: clblm_hip_handle = -1
   ! the reference's SAVEd JRAD (clblm_ODLAY.f90:182-183): starts at the control value
   integer(c_int32_t), save :: clblm_hip_jrad = 1
   ! arithmetic of the ODLAY drop-in: CLBLM_FP64 (the reference's to 1e-10) or CLBLM_FP32C
   integer(c_int), save :: clblm_hip_precision = CLBLM_FP64

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
      if (c_associated(clblm_hip_ctx)) return
      call clblm_hip_error(c_ctx_create(dev, clblm_hip_ctx), 'clblm_hip_init')
   end subroutine

   logical function clblm_hip_have_lines()
      clblm_hip_have_lines = clblm_hip_handle >= 0
   end function

   ! Upload the line list once (replaces the TAPE3 reads of RDLIN and RDLNFL).
   subroutine clblm_hip_lines(nu, s, a_air, a_self, e_lo, tmpalf, shift, mol, iflg)
      real(c_double), intent(in) :: nu(:)
      real(c_float), intent(in) :: s(:), a_air(:), a_self(:), e_lo(:), tmpalf(:), shift(:)
      integer(c_int32_t), intent(in) :: mol(:), iflg(:)
      call clblm_hip_init()
      call clblm_hip_error(c_upload(clblm_hip_ctx, int(size(nu), c_int64_t), nu, s, a_air, a_self, e_lo, tmpalf, &
                                    shift, mol, iflg, clblm_hip_handle), 'clblm_hip_lines')
   end subroutine

   ! Read and upload a TAPE3 line file (its panels overlapping [vlo, vhi]; all of them
   ! when vlo > vhi), the species-broadening fields with it.
   subroutine clblm_hip_lines_file(path, vlo, vhi)
      character(*), intent(in) :: path
      real(c_double), intent(in), optional :: vlo, vhi
      real(c_double) :: a, b
      integer(c_int64_t) :: n
      a = 1.; b = 0.
      if (present(vlo)) a = vlo
      if (present(vhi)) b = vhi
      call clblm_hip_init()
      call clblm_hip_error(c_load_tape3(clblm_hip_ctx, trim(path)//c_null_char, a, b, clblm_hip_handle, n), &
                           'clblm_hip_lines_file')
   end subroutine

   ! The MT_CKD water-vapour continuum table from the reference's netCDF file, with
   ! CONTNM's XSELF and XFRGN scale factors.
   subroutine clblm_hip_mtckd(path, xself, xfrgn)
      character(*), intent(in) :: path
      real(c_double), intent(in) :: xself, xfrgn
      call clblm_hip_init()
      call clblm_hip_error(c_mtckd_load_nc(clblm_hip_ctx, trim(path)//c_null_char, xself, xfrgn), 'clblm_hip_mtckd')
   end subroutine

   ! CONTNM's scale factors of the CO2, O3, O2, N2 and Rayleigh continua (layers with
   ! CLBLM_OD_CONTNM).
   subroutine clblm_hip_contnm_scales(xco2c, xo3cn, xo2cn, xn2cn, xrayl)
      real(c_double), intent(in) :: xco2c, xo3cn, xo2cn, xn2cn, xrayl
      call clblm_hip_init()
      call clblm_hip_error(c_contnm_scales(clblm_hip_ctx, xco2c, xo3cn, xo2cn, xn2cn, xrayl), &
                           'clblm_hip_contnm_scales')
   end subroutine

   ! One layer: molecule numbers mols(1:n) with columns w(1:n), p (hPa), T (K), Wtot;
   ! flags: CLBLM_OD_SPECIES_BROAD, CLBLM_OD_MTCKD, CLBLM_OD_CONTNM.
   subroutine clblm_hip_odlay_plain(mols, w, pave, tave, wtot, v1, v2, dvnormal, dvnarrow, sample, dvout, &
                                    precision, nsamp, od, dv_used, continuum, flags)
      integer, intent(in) :: mols(:)
      real(c_double), intent(in) :: w(:), pave, tave, wtot, v1, v2, dvnormal, dvnarrow, sample, dvout
      integer(c_int), intent(in) :: precision
      integer, intent(out) :: nsamp
      real(c_double), allocatable, target, intent(out) :: od(:)
      real(c_double), intent(out) :: dv_used
      real(c_double), intent(in), optional, target :: continuum(:)
      integer(c_int32_t), intent(in), optional :: flags
      type(clblm_layer_t) :: lay
      type(clblm_od_params_t) :: par
      integer(c_int32_t) :: jr
      integer(c_int64_t) :: ns
      integer :: k
      if (clblm_hip_handle < 0) stop '--- clblm_hip_odlay_plain(): no line list (clblm_hip_lines or _lines_file)'
      lay%pave = pave; lay%tave = tave; lay%wtot = wtot; lay%zbot = 0; lay%ztop = 0
      lay%w = 0; lay%sel = 0
      do k = 1, size(mols)
         lay%w(mols(k)) = w(k)
         lay%sel(mols(k)) = 1
      end do
      par%v1 = v1; par%v2 = v2; par%dvnormal = dvnormal; par%dvnarrow = dvnarrow; par%sample = sample
      par%dvout = dvout; par%precision = precision; par%contnm = 0; par%continuum = c_null_ptr
      par%nptabs = 0; par%flags = 0
      if (present(flags)) par%flags = flags
      if (present(continuum)) then
         par%contnm = 1
         par%continuum = c_loc(continuum)
         par%nptabs = size(continuum)
      end if
      jr = clblm_hip_jrad
      call clblm_hip_error(c_odlay(clblm_hip_ctx, clblm_hip_handle, lay, par, jr, ns, dv_used, c_null_ptr), &
                           'clblm_hip_odlay')
      allocate (od(ns))
      jr = clblm_hip_jrad
      call clblm_hip_error(c_odlay(clblm_hip_ctx, clblm_hip_handle, lay, par, jr, ns, dv_used, c_loc(od)), &
                           'clblm_hip_odlay')
      clblm_hip_jrad = jr
      nsamp = int(ns)
   end subroutine

   ! Device arrays for the batched entry points.
   type(c_ptr) function clblm_hip_dev_alloc(bytes)
      integer(c_int64_t), intent(in) :: bytes
      call clblm_hip_init()
      call clblm_hip_error(c_dev_alloc(clblm_hip_ctx, bytes, clblm_hip_dev_alloc), 'clblm_hip_dev_alloc')
   end function

   subroutine clblm_hip_dev_free(d)
      type(c_ptr), intent(inout) :: d
      call clblm_hip_error(c_dev_free(clblm_hip_ctx, d), 'clblm_hip_dev_free')
      d = c_null_ptr
   end subroutine

   subroutine clblm_hip_dev_put(d, h)
      type(c_ptr), intent(in) :: d
      real(c_double), intent(in), target :: h(:)
      call clblm_hip_error(c_dev_put(clblm_hip_ctx, d, c_loc(h), int(8 * size(h), c_int64_t)), 'clblm_hip_dev_put')
   end subroutine

   subroutine clblm_hip_dev_get(h, d)
      real(c_double), intent(inout), target :: h(:)
      type(c_ptr), intent(in) :: d
      call clblm_hip_error(c_dev_get(clblm_hip_ctx, c_loc(h), d, int(8 * size(h), c_int64_t)), 'clblm_hip_dev_get')
   end subroutine

   subroutine clblm_hip_finalize()
      integer(c_int) :: rc
      if (c_associated(clblm_hip_ctx)) rc = c_ctx_destroy(clblm_hip_ctx)
      clblm_hip_ctx = c_null_ptr
      clblm_hip_handle = -1
   end subroutine
end module
