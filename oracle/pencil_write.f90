! Fortran restatement of the reference's pencil checkpoint writer, TEST
! INFRASTRUCTURE ONLY (a checker of psdns_outpen and of oracle/pencil.py).
!
! The records are written by the Fortran runtime's own sequential
! unformatted I/O with big-endian conversion, as the reference's build does
! (PSDNS/makefile:251 -fconvert=big-endian), following the loop of
! PSDNS/homogeneous/outpen_write.f:68-76 for one task (xist = 1,
! xisz = n/2, zjst = 1, zjsz = n), with buf filled from the cylinder and zero
! off it (PSDNS/kxcomm1.f:586-612).
!
! Input (stream, native byte order): n, num_al, nv, file_prec (int32), then
! col_x(num_al), col_z(num_al) (int32, 0-based), then uny(n, num_al, nv)
! as complex(8).  Usage: pencil_write <input> <output>
program pencil_write
  implicit none
  integer :: n, na, nv, fprec, i, a, x, y, z
  integer, allocatable :: colx(:), colz(:)
  complex(8), allocatable :: uny(:,:,:), buf(:,:,:)
  complex(4), allocatable :: bufs(:,:,:)
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (10, file=trim(fin), access='stream', form='unformatted', status='old')
  read (10) n, na, nv, fprec
  allocate (colx(na), colz(na), uny(n, na, nv), buf(n/2, n, n), bufs(n/2, n, n))
  read (10) colx, colz, uny
  close (10)

  open (21, file=trim(fout), form='unformatted', convert='big_endian', status='replace')
  do i = 1, nv
    buf = (0.d0, 0.d0)
    do a = 1, na
      x = colx(a) + 1
      z = colz(a) + 1
      do y = 1, n
        buf(x, y, z) = uny(y, a, i)
      end do
    end do
    write (21) i, n, n, n
    write (21) 1, n/2, 1, n, 1, n
    if (fprec == 8) then
      do z = 1, n
        if (z == n/2 + 1) cycle
        write (21) ((buf(x, y, z), x = 1, n/2), y = 1, n/2)
        write (21) ((buf(x, y, z), x = 1, n/2), y = n/2 + 2, n)
      end do
    else
      bufs = cmplx(buf, kind=4)
      do z = 1, n
        if (z == n/2 + 1) cycle
        write (21) ((bufs(x, y, z), x = 1, n/2), y = 1, n/2)
        write (21) ((bufs(x, y, z), x = 1, n/2), y = n/2 + 2, n)
      end do
    end if
  end do
  close (21)
end program pencil_write
