! Fortran caller of the batched entry points through ISO_C_BINDING: reads a line list,
! layers and a spectral block from a binary file written by the test, computes every
! layer's OD on one grid with c_od_batch (device arrays from clblm_hip_dev_alloc), then
! the radiance and its per-layer OD derivatives with c_rt_jac, and writes them back.
program test_jac_f
   use iso_c_binding
   use clblm_hip
   implicit none
   integer :: u, k, l
   integer(c_int32_t) :: n, nlay, lin, jr
   real(c_double), allocatable :: nu(:), dvs(:), tlev(:), tlay(:), up(:), lo(:), rad(:), dtau(:), hw(:)
   real(c_float), allocatable :: s(:), a_air(:), a_self(:), e_lo(:), tmpalf(:), shift(:)
   integer(c_int32_t), allocatable :: mol(:), iflg(:)
   type(clblm_layer_t), allocatable :: lay(:)
   type(clblm_od_params_t), allocatable :: par(:)
   integer(c_int64_t), allocatable :: off(:)
   real(c_double) :: v1, v2, dvout, emis, tskin, sec
   integer(c_int64_t) :: npts
   type(c_ptr) :: d_od, d_tlay, d_tlev, d_up, d_lo, d_rad, d_dtau, d_dtlev, d_dts, d_de
   character(256) :: fin, fout
   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
   read (u) n, nlay
   allocate (nu(n), s(n), a_air(n), a_self(n), e_lo(n), tmpalf(n), shift(n), mol(n), iflg(n))
   read (u) nu, s, a_air, a_self, e_lo, tmpalf, shift, mol, iflg
   allocate (lay(nlay), par(nlay), off(nlay + 1), dvs(nlay), tlev(nlay + 1), tlay(nlay), up(nlay), lo(nlay))
   allocate (hw(10))
   do l = 1, nlay
      read (u) hw
      lay(l)%pave = hw(1); lay(l)%tave = hw(2); lay(l)%wtot = hw(3); lay(l)%zbot = 0; lay(l)%ztop = 0
      lay(l)%w = 0; lay(l)%sel = 0
      do k = 1, 7
         lay(l)%w(k) = hw(3 + k)
         lay(l)%sel(k) = 1
      end do
      tlay(l) = hw(2)
   end do
   read (u) v1, v2, dvout, dvs, tlev, up, lo, emis, tskin, sec, lin
   close (u)
   call clblm_hip_lines(nu, s, a_air, a_self, e_lo, tmpalf, shift, mol, iflg)
   do l = 1, nlay
      par(l)%v1 = v1; par(l)%v2 = v2; par(l)%dvnormal = dvs(l); par(l)%dvnarrow = dvs(l); par(l)%sample = 4.
      par(l)%dvout = dvout; par(l)%precision = CLBLM_FP64; par(l)%contnm = 0; par(l)%continuum = c_null_ptr
      par(l)%nptabs = 0; par(l)%flags = 0
   end do
   jr = 1
   call clblm_hip_error(c_od_batch(clblm_hip_ctx, clblm_hip_handle, nlay, lay, par, jr, off, c_null_ptr, &
                                   c_null_ptr), 'od_batch sizing')
   npts = off(2) - off(1)
   d_od = clblm_hip_dev_alloc(8 * off(nlay + 1))
   jr = 1
   call clblm_hip_error(c_od_batch(clblm_hip_ctx, clblm_hip_handle, nlay, lay, par, jr, off, d_od, c_null_ptr), &
                        'od_batch')
   d_tlay = clblm_hip_dev_alloc(8_c_int64_t * nlay)
   d_tlev = clblm_hip_dev_alloc(8_c_int64_t * (nlay + 1))
   d_up = clblm_hip_dev_alloc(8_c_int64_t * nlay)
   d_lo = clblm_hip_dev_alloc(8_c_int64_t * nlay)
   call clblm_hip_dev_put(d_tlay, tlay)
   call clblm_hip_dev_put(d_tlev, tlev)
   call clblm_hip_dev_put(d_up, up)
   call clblm_hip_dev_put(d_lo, lo)
   d_rad = clblm_hip_dev_alloc(8 * npts)
   d_dtau = clblm_hip_dev_alloc(8 * npts * nlay)
   d_dtlev = clblm_hip_dev_alloc(8 * npts * (nlay + 1))
   d_dts = clblm_hip_dev_alloc(8 * npts)
   d_de = clblm_hip_dev_alloc(8 * npts)
   call clblm_hip_error(c_rt_jac(clblm_hip_ctx, int(nlay, c_int), npts, v1, dvout, d_od, d_tlay, d_tlev, d_up, d_lo, &
                                 emis, tskin, int(lin, c_int), sec, d_rad, d_dtau, d_dtlev, d_dts, d_de, c_null_ptr), &
                        'rt_jac')
   allocate (rad(npts), dtau(npts * nlay))
   call clblm_hip_dev_get(rad, d_rad)
   call clblm_hip_dev_get(dtau, d_dtau)
   open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
   write (u) npts
   write (u) rad, dtau
   close (u)
   call clblm_hip_dev_free(d_od)
   call clblm_hip_dev_free(d_rad)
   call clblm_hip_dev_free(d_dtau)
   call clblm_hip_finalize()
end program
