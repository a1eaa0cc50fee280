! Fortran caller of the HIP library through ISO_C_BINDING: reads a line list and
! one layer from a binary file written by the test, computes the layer OD with
! clblm_hip_odlay_plain and writes it back.
program test_odlay_f
   use iso_c_binding
   use clblm_hip
   implicit none
   integer :: u, n, nmol, nsamp, k
   real(c_double), allocatable :: nu(:), od(:), w(:)
   real(c_float), allocatable :: s(:), a_air(:), a_self(:), e_lo(:), tmpalf(:), shift(:)
   integer(c_int32_t), allocatable :: mol(:), iflg(:)
   integer, allocatable :: mols(:)
   real(c_double) :: hdr(9), dvu
   character(256) :: fin, fout
   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
   read (u) n, nmol
   allocate (nu(n), s(n), a_air(n), a_self(n), e_lo(n), tmpalf(n), shift(n), mol(n), iflg(n), mols(nmol), w(nmol))
   read (u) nu, s, a_air, a_self, e_lo, tmpalf, shift, mol, iflg
   read (u) mols, w, hdr
   close (u)
   call clblm_hip_lines(nu, s, a_air, a_self, e_lo, tmpalf, shift, mol, iflg)
   ! hdr: pave, tave, wtot, v1, v2, dvnormal, dvnarrow, sample, dvout
   call clblm_hip_odlay_plain(mols, w, hdr(1), hdr(2), hdr(3), hdr(4), hdr(5), hdr(6), hdr(7), hdr(8), hdr(9), &
                              CLBLM_FP64, nsamp, od, dvu)
   open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
   write (u) nsamp, dvu, clblm_hip_jrad
   write (u) (od(k), k=1, nsamp)
   close (u)
   call clblm_hip_finalize()
end program
