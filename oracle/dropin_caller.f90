! A caller of the reference's generic ODLAY written as the reference's own callers are
! (layerMerge_*, drv_LevelTemp: clblm_noScattRT.f90, clblm_noScattJac.f90), built
! against the reference's Module_ConstParam and Module_Config and the drop-in
! Module_ODLAY (clblm_amd/fortran/module_odlay_hip.f90) in place of clblm_ODLAY.f90.
! It sets ioFiles%lineFile and the OD control record, calls ODLAY for a few layers and
! writes the optical depths.  Test infrastructure: it links the reference's modules.
program dropin_caller
   use Module_ConstParam, only: r8 => kind_r8
   use Module_Config, only: CLBLM_OD_Ctrl, ioFiles
   use Module_ODLAY, only: ODLAY
   implicit none
   type(CLBLM_OD_Ctrl) :: odCtrl
   integer :: u, nlay, l, nSamp, k, cont, brd
   character(256) :: fin, fout, tape3
   character(len=6) :: molID(7) = ['H2O   ', 'CO2   ', 'O3    ', 'N2O   ', 'CO    ', 'CH4   ', 'O2    ']
   real :: Pav, Tav, W(7), Wtot, DVn, DVout
   real(r8) :: V1, V2
   real, allocatable :: monoOD(:)
   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
   read (u) tape3
   read (u) nlay, cont, brd, V1, V2
   ioFiles%lineFile = trim(tape3)
   odCtrl%contnmOff = cont == 0
   odCtrl%speciesBroad = brd /= 0
   odCtrl%JRAD = 1
   open (newunit=k, file=trim(fout), access='stream', form='unformatted', status='replace')
   do l = 1, nlay
      read (u) Pav, Tav, Wtot, W, DVn
      DVout = 0.
      call ODLAY(7, molID, Pav, Tav, W, Wtot, 0., 1., odCtrl, V1, V2, DVn, DVn, 4., 0.04, 0, l, nSamp, monoOD, &
                 DVout=DVout)
      write (k) nSamp, DVout
      write (k) monoOD(1:nSamp)
   end do
   close (k)
   close (u)
end program
