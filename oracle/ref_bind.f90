! C-callable entry points into the reference's own Fortran modules, compiled from
! the reference sources where they lie (see the `ref` target of oracle/Makefile).
! Test infrastructure only: used to pin the oracle's tables and helper routines
! against the reference bit for bit.  Nothing here is part of the product.
module ref_bind
   use iso_c_binding
   use Module_ConstParam,      only: r8 => kind_r8, clblmMolNames, MXMOL, MXISOTPL
   use Module_ODLAY_CommonSub, only: MOLEC, RADFN, RADFNI, AVRAT, CGAUSS, CF1, CF2, CF3, CER
   use Module_TIPS,            only: TIPS_2003
   use Module_Spectrum,        only: XINT, interpSpectrum, CLBLM_Spectrum, CLBLM_Spectrum_init, interpToFinestGrid
   use Module_EMLAY,           only: planck, planckFnDrv, rlin, layerEmis
   use Module_FFT,             only: clblm_FFTSCN, Intpdr
   use Module_ScanFilter,      only: clblm_SCANFN, clblm_FLTRFN
   use Module_LineData,        only: INPUT_HEADER, INPUT_BLOCK, NLINEREC, MXBRDMOL, MXBRDMOL3
   implicit none
   external :: BUFIN_sgl
contains

   ! A TAPE3 line file read with the reference's own record types (INPUT_HEADER and
   ! INPUT_BLOCK, clblm_LineData.f90:52-70) and record reader (BUFIN_sgl,
   ! util_linux_intel.f90:121-150), in the order RDLIN and RDLNFL read it
   ! (clblm_ODLAY.f90:2626-2658, clblm_ODLAY_LineF4.f90:1408-1418): the header record
   ! (as lnfilhd_4 reads it, clblm_ODLAY_LineF4.f90:1302-1309), the negative-E'' record
   ! when HLINID(7) ends in '^', then per panel its 6-word header and its NWDS-word
   ! block, until the end of the file.  Every panel is returned, unmodified (RDLIN's
   ! O2/N2 width correction is applied after the read).  Panel q (1-based) holds
   ! entries 250*(q-1)+1 .. 250*(q-1)+nrec(q) of the per-entry arrays.
   subroutine ref_tape3(path, cap, npanel, nrec, nwds, vmin, vmax, vnu, sp, alfa, epp, mol, hwhm, tmpalf, &
                        pshift, iflg, brdflg, brddat, sdep, hdr_i, hdr_f) bind(C, name="ref_tape3")
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: cap
      integer(c_int), intent(out) :: npanel, nrec(cap), nwds(cap), mol(NLINEREC, cap), iflg(NLINEREC, cap)
      integer(c_int), intent(out) :: brdflg(MXBRDMOL, NLINEREC, cap), hdr_i(5)
      real(c_double), intent(out) :: vmin(cap), vmax(cap), vnu(NLINEREC, cap)
      real(c_float), intent(out) :: sp(NLINEREC, cap), alfa(NLINEREC, cap), epp(NLINEREC, cap)
      real(c_float), intent(out) :: hwhm(NLINEREC, cap), tmpalf(NLINEREC, cap), pshift(NLINEREC, cap)
      real(c_float), intent(out) :: brddat(MXBRDMOL3, NLINEREC, cap), sdep(NLINEREC, cap), hdr_f(2)
      type(INPUT_HEADER) :: pnl
      type(INPUT_BLOCK) :: buf
      character(8) :: hlinid(10), bmolid(64), hid1(2)
      integer*4 :: molcnt(64), mcntlc(64), mcntnl(64), linmol, lincnt, ilinlc, ilinnl, irec, irectl
      integer*4 :: nneg(64), nres(64)
      real*4 :: sumstr(64), flinlo, flinhi, xspace(4096)
      character(len=4096) :: fname
      integer :: lun, leof, n, q
      n = 0
      do while (path(n + 1) /= c_null_char .and. n < len(fname))
         n = n + 1
         fname(n:n) = path(n)
      end do
      open(newunit=lun, file=fname(1:n), status='OLD', form='UNFORMATTED')
      read(lun) hlinid, bmolid, molcnt, mcntlc, mcntnl, sumstr, linmol, flinlo, flinhi, lincnt, ilinlc, &
                ilinnl, irec, irectl, hid1
      hdr_i = [linmol, lincnt, ilinlc, 0, 0]
      hdr_f = [flinlo, flinhi]
      if (hlinid(7)(8:8) == '^') then
         hdr_i(4) = 1
         read(lun) nneg, nres, xspace
      end if
      if (hlinid(10)(1:1) == 'I') hdr_i(5) = 1
      npanel = 0
      do q = 1, cap
         call BUFIN_sgl(lun, leof, pnl, 6)
         if (leof == 0) exit
         call BUFIN_sgl(lun, leof, buf, pnl%NWDS)
         npanel = q
         nrec(q) = pnl%NREC; nwds(q) = pnl%NWDS; vmin(q) = pnl%VMIN; vmax(q) = pnl%VMAX
         vnu(:, q) = buf%VNU; sp(:, q) = buf%SP; alfa(:, q) = buf%ALFA; epp(:, q) = buf%EPP
         mol(:, q) = buf%MOL; hwhm(:, q) = buf%HWHM; tmpalf(:, q) = buf%TMPALF; pshift(:, q) = buf%PSHIFT
         iflg(:, q) = buf%IFLG; brdflg(:, :, q) = buf%BRD_MOL_FLG_IN; brddat(:, :, q) = buf%BRD_MOL_DAT
         sdep(:, q) = buf%SPEED_DEP
      end do
      close(lun)
   end subroutine

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

   ! clblm_FLTRFN (clblm_ScanFilter.f90:1703-1860): the spectrum sp(1:np) on v1 + (i-1) dv
   ! through the filter xf(1:nptf) from v1f at dvf (dvf < 0: v1f is the centre)
   subroutine ref_fltrfn(np, sp, v1, dv, v1f, dvf, nptf, xf, out) bind(C, name="ref_fltrfn")
      integer(c_int), value :: np, nptf
      real(c_double), value :: v1, dv, v1f, dvf
      real(c_double), intent(in) :: sp(np), xf(nptf)
      real(c_double), intent(out) :: out
      type(CLBLM_Spectrum) :: s
      real(r8) :: a, b
      real :: d, e, f
      real, allocatable :: x(:)
      a = v1; d = dv
      call CLBLM_Spectrum_init(s, a, d, np)
      s%spect(s%indV1:s%indV2) = sp
      b = v1f; e = dvf
      allocate(x(nptf))
      x = xf
      f = 0.
      call clblm_FLTRFN(s, b, e, nptf, x, f)
      out = f
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

   ! layerEmis (clblm_EMLAY.f90:104-178): emission of a layer of optical depth od(1:np)
   ! on v1 + (i-1) dv from its Ta side, transmittance exp(-od), linInTau lin.
   subroutine ref_layer_emis(np, od, v1, dv, ta, tb, tave, lin, out) bind(C, name="ref_layer_emis")
      integer(c_int), value :: np, lin
      real(c_double), value :: v1, dv, ta, tb, tave
      real(c_double), intent(in) :: od(np)
      real(c_double), intent(out) :: out(np)
      type(CLBLM_Spectrum) :: em, tx, sod
      real(r8) :: a
      real :: d
      a = v1; d = dv
      call CLBLM_Spectrum_init(sod, a, d, np)
      call CLBLM_Spectrum_init(tx, a, d, np)
      sod%spect(sod%indV1:sod%indV2) = od
      tx%spect(tx%indV1:tx%indV2) = exp(-od)
      call layerEmis(em, ta, tb, tave, tx, sod, lin)
      out = em%spect(em%indV1:em%indV1 + np - 1)
   end subroutine

   ! interpToFinestGrid (clblm_Spectrum.f90:799-866) of nspec spectra from v1: spectrum k
   ! holds nlim(k) values at data(off(k)+1 ..) on dv(k).  On return spectrum k is on the
   ! finest grid: nlim_out(k) values at out(k*cap+1 ..), dv_out(k).
   subroutine ref_interp_to_finest(nspec, v1, dv, nlim, off, data, cap, out, nlim_out, dv_out) &
         bind(C, name="ref_interp_to_finest")
      integer(c_int), value :: nspec, cap
      real(c_double), value :: v1
      real(c_double), intent(in) :: dv(nspec), data(*)
      integer(c_int64_t), intent(in) :: nlim(nspec), off(nspec)
      real(c_double), intent(out) :: out(cap, nspec), dv_out(nspec)
      integer(c_int64_t), intent(out) :: nlim_out(nspec)
      type(CLBLM_Spectrum), allocatable :: sp(:)
      real(r8) :: a
      real :: d
      integer :: k, n
      allocate(sp(nspec))
      a = v1
      do k = 1, nspec
         d = dv(k)
         n = int(nlim(k))
         call CLBLM_Spectrum_init(sp(k), a, d, n)
         sp(k)%spect(sp(k)%indV1:sp(k)%indV2) = data(off(k) + 1:off(k) + n)
      end do
      call interpToFinestGrid(sp)
      do k = 1, nspec
         n = min(sp(k)%NLIM, cap)
         nlim_out(k) = sp(k)%NLIM
         dv_out(k) = sp(k)%DV
         out(1:n, k) = sp(k)%spect(sp(k)%indV1:sp(k)%indV1 + n - 1)
      end do
   end subroutine

   subroutine ref_rlin(od, hlf, dhlf) bind(C, name="ref_rlin")
      real(c_double), value :: od
      real(c_double), intent(out) :: hlf, dhlf
      call rlin(od, hlf, dhlf)
   end subroutine
end module
