! Writes sample lines with the edit descriptors of the reference's text
! outputs (qm_eng.d: Sources/pwlda.f90:2598; qm_eig.d: Sources/fermi.f90:
! 1924-1937; qm_fer.d: 1952) so that the Python writers can be checked
! against flang's own formatting.  Built and run by make_golden.py.
program fortran_formats
implicit none
real*8 :: e(16), eig(4), occ(4)
integer :: i
e = (/ -31.4159265358979d0, -31.4159265d0, 12.5d0, -3.25d-3, 7.125d0, -8.0d0, 1.0d-12,  &
&      0.d0, 2.5d10, -1.5d-7, 3.d0, 4.d0, 5.d0, 6.d0, 7.d0, 8.d0 /)
write(*,'(2i6,2es18.10,20es14.6)') 17, 3, e(1), e(2), e(3:16)
write(*,'(10i7)') 17, 3, 4
eig = (/ -0.512345678d0, -0.25d0, 0.1d0, 1.75d0 /)
occ = (/ 2.d0, 1.5d0, 0.d0, -0.25d0 /)
do i = 1, 4
  if (occ(i) >= 0.d0) then
    write(*,'(i6,es13.5,f6.3)') i, eig(i), occ(i)
  else
    write(*,'(i6,es13.5,f7.3)') i, eig(i), occ(i)
  end if
end do
write(*,'(2i7,2es13.5)') 17, 3, -0.123456789d0
end program
