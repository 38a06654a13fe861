! Writes sample lines with the edit descriptors of the reference's text
! outputs (qm_eng.d: Sources/pwlda.f90:2598; qm_eig.d: Sources/fermi.f90:
! 1924-1937; qm_fer.d: 1952; qm_ion.d and qm_frc.d: Sources/datadump.f90:
! 235-266) so that the Python writers can be checked against flang's own
! formatting.  Built and run by make_golden.py.
program fortran_formats
implicit none
real*8 :: e(16), eig(4), occ(4), frc(3, 5), vatm(3, 5), atconfx
integer :: i, ia, nhk(2)
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
! qm_frc.d block: forces scaled by max |f| / 9.999
nhk = (/ 3, 2 /)
frc = reshape((/ 0.0123d0, -0.5d0, 0.25d0, 1.d-9, -0.031d0, 0.0d0, 0.49999d0, -0.125d0,  &
&     0.3333333d0, -0.2d0, 0.07d0, -0.0049d0, 0.1d0, 0.45d0, -0.0001d0 /), (/ 3, 5 /))
atconfx = 0.d0
do i = 1, 3
do ia = 1, 5
   atconfx = max( atconfx, abs(frc(i,ia)) )
end do
end do
atconfx = atconfx / 9.999d0
write(*,'(100i7)') 17, 2, (nhk(i), i = 1, 2)
write(*,'(es14.7)') atconfx
if( abs(atconfx) < 1.d-30 ) atconfx = 1.d0
write(*,'(9f8.5)') ( ( frc(i,ia)/atconfx, i = 1, 3 ), ia = 1, 5 )
! qm_ion.d block: scaled coordinates / 0.1, capped at 9.99999
vatm = reshape((/ 0.d0, 0.25d0, 0.5d0, 0.999999d0, 0.1234567d0, 0.75d0, 1.5d0, 0.0625d0,  &
&     0.3d0, 0.87654321d0, 0.01d0, 0.999d0, 0.4d0, 0.6d0, 0.05d0 /), (/ 3, 5 /))
atconfx = 0.1d0
write(*,'(100i7)') 17, 2, (nhk(i), i = 1, 2)
write(*,'(es14.7)') atconfx
write(*,'(9f8.5)') ( ( min(vatm(i,ia)/atconfx,9.99999d0), i = 1, 3 ), ia = 1, 5 )
end program
