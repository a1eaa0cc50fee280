! Drop-in Module_ODLAY: the reference's generic ODLAY (src/clblm_src/clblm_ODLAY.f90:12-18)
! with both of its forms, ODLAY_struct (:26-124, the form every caller uses) and
! ODLAY_nonStruct (:129-176), with the same dummy arguments, computed by the HIP
! library through clblm_hip.  Compile it with -cpp in place of clblm_ODLAY.f90,
! against the reference's modules; the callers (layerMerge_*, pathOD_up_dn_sun,
! drv_LevelTemp, drv_LevelConc) are unchanged.  ODLAY_struct needs the reference's
! Module_AtmPath and Module_DV, which come with its netCDF-dependent scene modules;
! -DCLBLM_HIP_NO_STRUCT leaves it out for builds without them.
! As the reference does (:276-302), the first call reads the line file named by
! ioFiles%lineFile (a TAPE3 file, line coupling and species-broadening fields
! included) unless a list was uploaded with clblm_hip_lines.  odCtrl%speciesBroad
! selects the species-by-species broadening.  With the continuum on
! (.not. odCtrl%contnmOff) CONTNM's continua are added: the MT_CKD water vapour (self and
! foreign, XSELF and XFRGN) from the reference's coefficient file, read on the first
! call, and CONTNM's other continua: CO2, O3, the O2 and N2 collision-induced and
! electronic bands and Rayleigh extinction (XCO2C, XO3CN, XO2CN, XN2CN, XRAYL).
! clblm_hip_precision selects FP64 (default) or FP32C.  Not supported: NLTE,
! cross-sections, lineOff.
module Module_ODLAY
   use iso_c_binding
   use clblm_hip
   implicit none
   private
   public :: ODLAY
   interface ODLAY
#ifndef CLBLM_HIP_NO_STRUCT
      module procedure ODLAY_struct
#endif
      module procedure ODLAY_nonStruct
   end interface
   logical, save :: firstPassJRAD = .true.
   logical, save :: mtckdLoaded = .false.
   ! the reference's MT_CKD coefficient file (mt_ckd_h2o_module.f90:16)
   character(*), parameter :: mtckdFile = 'clblm_data/spectroscopy/absco-ref_wv-mt-ckd.nc'
contains
#ifndef CLBLM_HIP_NO_STRUCT
   ! Unpacks the layer, grid and DV records as the reference's wrapper does and
   ! returns the OD as a CLBLM_Spectrum on the grid actually used (V1, DVout).
   subroutine ODLAY_struct(aLayer, odCtrl, spectGrid, dvCtrl, layOD, NLTE, layNLTEEmisFac, JacMolName)
      use Module_Config, only: CLBLM_OD_Ctrl, CLBLM_DV_Ctrl
      use Module_DV, only: CLBLM_SpectGrid
      use Module_Spectrum, only: CLBLM_Spectrum, CLBLM_Spectrum_init
      use Module_AtmPath, only: CLBLM_Layer
      type(CLBLM_Layer), intent(in) :: aLayer
      type(CLBLM_OD_Ctrl), intent(in) :: odCtrl
      type(CLBLM_SpectGrid), intent(in) :: spectGrid
      type(CLBLM_DV_Ctrl), intent(in) :: dvCtrl
      type(CLBLM_Spectrum), intent(out) :: layOD
      logical, optional, intent(in) :: NLTE
      type(CLBLM_Spectrum), optional, intent(out) :: layNLTEEmisFac
      character(*), optional, intent(in) :: JacMolName
      integer :: nSamp, L
      real, allocatable :: monoOD(:)
      real :: dvo
      if (present(NLTE)) then
         if (NLTE) stop '--- ODLAY(): NLTE is not supported by the HIP path'
      end if
      ! JacMolName only selects the continuum derivative of CONTNM's AJ branch
      ! (clblm_ODLAY.f90:438), which this path does not compute
      L = aLayer%layNo
      dvo = spectGrid%DVout
      call ODLAY_nonStruct(aLayer%nMol, aLayer%molID, aLayer%Pave, aLayer%Tave, aLayer%W, aLayer%Wtot, &
                           aLayer%Zbot, aLayer%Ztop, odCtrl, spectGrid%V1, spectGrid%V2, &
                           spectGrid%DVnormal(L), spectGrid%DVnarrow(L), dvCtrl%SAMPLE, dvCtrl%ALFAL0, &
                           0, L, nSamp, monoOD, DVout=dvo)
      call CLBLM_Spectrum_init(layOD, spectGrid%V1, dvo, nSamp)
      call move_alloc(monoOD, layOD%spect)
   end subroutine
#endif

   subroutine ODLAY_nonStruct(nMol, molID, Pav, Tav, W, Wtot, Zbot, Ztop, odCtrl, &
                              V1, V2, DVnormal, DVnarrow, &
                              SAMPLE, ALFAL0, IPATH, LayNo, &
                              nSamp, monoOD, NLTE, nlteEmisFac, DVout, ajMolName)
      use Module_ConstParam, only: r8 => kind_r8, molNum
      use Module_Config, only: CLBLM_OD_Ctrl, ioFiles
      integer, intent(in) :: nMol
      character(*), intent(in) :: molID(:)
      real, intent(in) :: Pav, Tav, W(:), Wtot, Zbot, Ztop
      type(CLBLM_OD_Ctrl), intent(in) :: odCtrl
      real(r8), intent(in) :: V1, V2
      real, intent(in) :: DVnormal, DVnarrow, SAMPLE, ALFAL0
      integer, intent(in) :: IPATH, layNo
      integer, intent(out) :: nSamp
      real, allocatable, intent(out) :: monoOD(:)
      logical, optional, intent(in) :: NLTE
      real, allocatable, optional, intent(out) :: nlteEmisFac(:)
      real, optional, intent(inout) :: DVout
      character(*), optional, intent(in) :: ajMolName
      integer :: mols(nMol), k
      integer(c_int32_t) :: flags
      real(c_double) :: dvo, dvu
      real(c_double), allocatable :: od(:)
      if (present(NLTE)) then
         if (NLTE) stop '--- ODLAY(): NLTE is not supported by the HIP path'
      end if
      if (odCtrl%lineOff) stop '--- ODLAY(): lineOff is not supported by the HIP path'
      if (.not. clblm_hip_have_lines()) call clblm_hip_lines_file(trim(ioFiles%lineFile))
      flags = 0
      if (odCtrl%speciesBroad) flags = ior(flags, CLBLM_OD_SPECIES_BROAD)
      if (.not. odCtrl%contnmOff) then
         if (.not. mtckdLoaded) then
            call clblm_hip_mtckd(mtckdFile, real(odCtrl%XSELF, c_double), real(odCtrl%XFRGN, c_double))
            call clblm_hip_contnm_scales(real(odCtrl%XCO2C, c_double), real(odCtrl%XO3CN, c_double), &
                                         real(odCtrl%XO2CN, c_double), real(odCtrl%XN2CN, c_double), &
                                         real(odCtrl%XRAYL, c_double))
            mtckdLoaded = .true.
         end if
         flags = ior(flags, ior(CLBLM_OD_MTCKD, CLBLM_OD_CONTNM))
      end if
      do k = 1, nMol
         mols(k) = molNum(molID(k))
      end do
      ! the first call takes JRAD from the control record, later calls keep the
      ! SAVEd value (clblm_ODLAY.f90:349-352)
      if (firstPassJRAD) then
         clblm_hip_jrad = odCtrl%JRAD
         firstPassJRAD = .false.
      end if
      dvo = 0.
      if (present(DVout)) then
         if (DVout > 0.) dvo = DVout
      end if
      call clblm_hip_odlay_plain(mols, real(W(1:nMol), c_double), real(Pav, c_double), real(Tav, c_double), &
                                 real(Wtot, c_double), real(V1, c_double), real(V2, c_double), &
                                 real(DVnormal, c_double), real(DVnarrow, c_double), real(SAMPLE, c_double), &
                                 dvo, clblm_hip_precision, nSamp, od, dvu, flags=flags)
      allocate (monoOD(nSamp))
      monoOD = od
      if (present(DVout)) DVout = dvu
   end subroutine
end module
