! C-callable entry points into the reference's own Fortran modules, compiled from
! the reference sources where they lie (see the `ref` target of oracle/Makefile).
! Test infrastructure only: used to pin the oracle's tables and helper routines
! against the reference bit for bit.  Nothing here is part of the product.
module ref_bind
   use iso_c_binding
   use Module_ConstParam,      only: r8 => kind_r8, clblmMolNames, MXMOL, MXISOTPL
   use Module_ODLAY_CommonSub, only: MOLEC, RADFN, RADFNI, AVRAT, CGAUSS, CF1, CF2, CF3, CER
   use Module_TIPS,            only: TIPS_2003
   use Module_Spectrum,        only: XINT, interpSpectrum, CLBLM_Spectrum, CLBLM_Spectrum_init
   use Module_EMLAY,           only: planck, planckFnDrv, rlin
   use Module_FFT,             only: clblm_FFTSCN, Intpdr
   use Module_ScanFilter,      only: clblm_SCANFN
   implicit none
contains

   ! clblm_SCANFN_array (clblm_ScanFilter.f90:33-67): spectral-domain scan of sp(1:np)
   ! on v1 + (i-1) dv with scanning function functid of half width hwhm onto
   ! v1out..v2out at dvout.
   subroutine ref_scanfn(np, sp, v1, v2, dv, v1out, v2out, dvout, functid, hwhm, cap, out, nout, &
                         v1o, v2o, dvo) bind(C, name="ref_scanfn")
      integer(c_int), value :: np, functid, cap
      real(c_double), value :: v1, v2, dv, v1out, v2out, dvout, hwhm
      real(c_double), intent(in) :: sp(np)
      real(c_double), intent(out) :: out(cap)
      integer(c_int), intent(out) :: nout
      real(c_double), intent(out) :: v1o, v2o, dvo
      real, allocatable :: d(:)
      real(r8) :: a, b, c1, c2
      real :: dvs, dvx, hw
      integer :: n
      a = v1; b = v2; c1 = v1out; c2 = v2out; dvs = dv; dvx = dvout; hw = hwhm
      n = 0
      call clblm_SCANFN(sp, a, b, dvs, np, d, c1, c2, dvx, n, functid, hw)
      nout = n
      v1o = c1; v2o = c2; dvo = dvx
      out(1:min(n, cap)) = d(1:min(n, cap))
   end subroutine

   ! Intpdr (clblm_FFT.f90:2392-2443): sp(1:np) on v1 + (i-1) dv onto vv1..vv2 at dvv
   subroutine ref_intpdr(np, sp, v1, dv, vv1, vv2, dvv, cap, out, nout, v1o) bind(C, name="ref_intpdr")
      integer(c_int), value :: np, cap
      real(c_double), value :: v1, dv, vv1, vv2, dvv
      real(c_double), intent(in) :: sp(np)
      real(c_double), intent(out) :: out(cap)
      integer(c_int), intent(out) :: nout
      real(c_double), intent(out) :: v1o
      type(CLBLM_Spectrum) :: s
      real(r8) :: a, b, c
      real :: d, e
      integer :: ierr, m
      c = v1; d = dv
      call CLBLM_Spectrum_init(s, c, d, np)
      s%spect(s%indV1:s%indV2) = sp
      a = vv1; b = vv2; e = dvv
      call Intpdr(s, a, b, e, ierr)
      nout = s%NLIM
      v1o = s%V1
      m = min(nout, cap)
      out(1:m) = s%spect(s%indV1:s%indV1 + m - 1)
   end subroutine

   ! clblm_FFTSCN_array (clblm_FFT.f90:308-349): scan of the spectrum sp(1:np) on
   ! v1 + (i-1) dv with scanning function functid of half width hwhm, limits
   ! v1out..v2out, output grid dvout (<= 0: the scan grid), no pre-boxcar.
   subroutine ref_fftscn(np, sp, v1, v2, dv, v1out, v2out, dvout, functid, hwhm, fill, cap, out, nout, &
                         v1o, v2o, dvo) bind(C, name="ref_fftscn")
      integer(c_int), value :: np, functid, cap
      real(c_double), value :: v1, v2, dv, v1out, v2out, dvout, hwhm, fill
      real(c_double), intent(in) :: sp(np)
      real(c_double), intent(out) :: out(cap)
      integer(c_int), intent(out) :: nout
      real(c_double), intent(out) :: v1o, v2o, dvo
      real, allocatable :: d(:)
      real(r8) :: a, b, c1, c2
      real :: dvx, dvs, hw, fl, parms(3)
      integer :: n
      a = v1; b = v2; c1 = v1out; c2 = v2out; dvx = dvout; dvs = dv; hw = hwhm; fl = fill
      parms = 0.
      n = 0
      call clblm_FFTSCN(sp, a, b, dvs, np, d, c1, c2, dvx, n, functid, hw, fl, -1.0, .false., parms)
      nout = n
      v1o = c1; v2o = c2; dvo = dvx
      out(1:min(n, cap)) = d(1:min(n, cap))
   end subroutine

   subroutine ref_fftscn2(np, sp, v1, v2, dv, v1out, v2out, dvout, functid, hwhm, fill, preboxhw, deconv, parm, &
                          cap, out, nout, v1o, v2o, dvo) bind(C, name="ref_fftscn2")
      integer(c_int), value :: np, functid, cap, deconv
      real(c_double), value :: v1, v2, dv, v1out, v2out, dvout, hwhm, fill, preboxhw
      real(c_double), intent(in) :: sp(np), parm(3)
      real(c_double), intent(out) :: out(cap)
      integer(c_int), intent(out) :: nout
      real(c_double), intent(out) :: v1o, v2o, dvo
      real, allocatable :: d(:)
      real(r8) :: a, b, c1, c2
      real :: dvx, dvs, hw, fl, pb, parms(3)
      integer :: n
      a = v1; b = v2; c1 = v1out; c2 = v2out; dvx = dvout; dvs = dv; hw = hwhm; fl = fill; pb = preboxhw
      parms = parm
      n = 0
      call clblm_FFTSCN(sp, a, b, dvs, np, d, c1, c2, dvx, n, functid, hw, fl, pb, deconv /= 0, parms)
      nout = n
      v1o = c1; v2o = c2; dvo = dvx
      out(1:min(n, cap)) = d(1:min(n, cap))
   end subroutine

   subroutine ref_zeta_tables(out) bind(C, name="ref_zeta_tables")
      real(c_double), intent(out) :: out(102, 6)
      out(:, 1) = AVRAT; out(:, 2) = CGAUSS; out(:, 3) = CF1
      out(:, 4) = CF2;   out(:, 5) = CF3;    out(:, 6) = CER
   end subroutine

   function ref_radfn(vi, xkt) bind(C, name="ref_radfn") result(r)
      real(c_double), value :: vi, xkt
      real(c_double) :: r
      real(r8) :: v
      v = vi
      r = RADFN(v, xkt)
   end function

   function ref_radfni(vi, dvi, xkt, vinew, rdel, rdlast) bind(C, name="ref_radfni") result(r)
      real(c_double), value :: vi, dvi, xkt
      real(c_double), intent(inout) :: vinew, rdlast
      real(c_double), intent(out) :: rdel
      real(c_double) :: r
      real(r8) :: v, vn
      v = vi; vn = vinew
      r = RADFNI(v, dvi, xkt, vn, rdel, rdlast)
      vinew = vn
   end function

   ! SCOR(mol, 1:10) from TIPS_2003 and ALFD1(mol, 1:10) from MOLEC at (p, T)
   subroutine ref_molec(mol, p, t, wk_mol, wtot, scor, alfd1, rhoslf) bind(C, name="ref_molec")
      integer(c_int), value :: mol
      real(c_double), value :: p, t, wk_mol, wtot
      real(c_double), intent(out) :: scor(10), alfd1(10), rhoslf
      real :: sc(MXMOL, MXISOTPL), rh(MXMOL), ad(MXMOL, MXISOTPL), wk(60)
      character(6) :: names(1)
      names(1) = clblmMolNames(mol)
      wk = 0.; wk(mol) = wk_mol
      sc = 0.; rh = 0.; ad = 0.
      call MOLEC(1, sc, rh, ad, p, t, wk, wtot, names)
      call MOLEC(2, sc, rh, ad, p, t, wk, wtot, names)
      scor = sc(mol, 1:10); alfd1 = ad(mol, 1:10); rhoslf = rh(mol)
   end subroutine

   subroutine ref_xint(v1a, v2a, dva, na, a, vft, dvr3, nr3, r3, n1r3, n2r3) bind(C, name="ref_xint")
      real(c_double), value :: v1a, v2a, dva, vft, dvr3
      integer(c_int), value :: na, nr3, n1r3, n2r3
      real(c_double), intent(in) :: a(na)
      real(c_double), intent(inout) :: r3(nr3)
      real(r8) :: v1, v2, vf
      v1 = v1a; v2 = v2a; vf = vft
      call XINT(v1, v2, dva, a, 1.0, vf, dvr3, r3, n1r3, n2r3)
   end subroutine

   subroutine ref_interp(nin, spin, dvin, v1in, nout, spout, dvout, v1out) bind(C, name="ref_interp")
      integer(c_int), value :: nin, nout
      real(c_double), value :: dvin, v1in, dvout, v1out
      real(c_double), intent(in) :: spin(nin)
      real(c_double), intent(out) :: spout(nout)
      real(r8) :: a, b
      real :: x(nin), y(nout)
      a = v1in; b = v1out
      x = spin
      call interpSpectrum(x, dvin, a, nin, y, dvout, b, nout)
      spout = y
   end subroutine

   function ref_planck(v, t) bind(C, name="ref_planck") result(r)
      real(c_double), value :: v, t
      real(c_double) :: r
      real(r8) :: vv
      real :: xkt
      vv = v; xkt = t / 1.4387752
      r = planck(vv, xkt)
   end function

   subroutine ref_planck_drv(v, t, b, dbdt) bind(C, name="ref_planck_drv")
      real(c_double), value :: v, t
      real(c_double), intent(out) :: b, dbdt
      real(r8) :: vv
      vv = v
      call planckFnDrv(vv, t, b, dbdt)
   end subroutine

   subroutine ref_rlin(od, hlf, dhlf) bind(C, name="ref_rlin")
      real(c_double), value :: od
      real(c_double), intent(out) :: hlf, dhlf
      call rlin(od, hlf, dhlf)
   end subroutine
end module
