!! ORACLE — test infrastructure only (never linked into the product library).
!!
!! Driver for the reference Femlib element routines compiled from their own
!! sources under the reference tree (see oracle/Makefile).  It reads one
!! command per line on stdin and prints the routine outputs with 17
!! significant digits, one value per line, so the Python tests can compare
!! the C restatement in oracle/elements.c against the reference itself.
!!
!! Commands: the command word on one line, its numbers on the next line:
!!   hexa31 E nu x1..x8 y1..y8 z1..z8           -> EK(24,24) column-major
!!   andes4 E nu t x1..x4 y1..y4 z1..z4         -> trel(3,4), Kloc(24,24), P(24,24)
!!   fqs35  rho x1..x4 y1..y4 z1..z4 t1..t4     -> EM(24)
!!   quit
program ref_driver
  use KindModule       , only : dp
  use IsoMatModule     , only : isoMat2D, isoMat3D
  use Andes4ShellModule, only : Andes4shell_stiffmat
  use ManipMatrixModule, only : trans3P
  use PmatModule       , only : pMatStiff
  implicit none
  character(len=16) :: cmd
  real(dp) :: E, nu, t, rho, X(8), Y(8), Z(8), EK(24,24), A1(24,9), A2(9,9)
  real(dp) :: C6(6,6), C3(3,3), Cmat(6,6), trel(3,4), Tel(3,3), tmp(3)
  real(dp) :: coorG(3,4), Xl(4), Yl(4), Zl(4), Kloc(24,24), P(24,24), EMV(24), TH(4)
  integer  :: ierr, i, j, ios
  do
     read(*,*,iostat=ios) cmd
     if (ios /= 0) exit
     select case (trim(cmd))
     case ('hexa31')
        read(*,*) E, nu, X, Y, Z
        call isoMat3D (E,nu,C6)
        call HEXA31 (EK,A1,A2,C6,X,Y,Z,3,0,1,6,0,ierr)
        write(*,'(I4)') ierr
        write(*,'(ES25.17)') EK
     case ('andes4')
        read(*,*) E, nu, t, X(1:4), Y(1:4), Z(1:4)
        !! Same sequence as ADD_ANDES4 (inadd.f90:1452-1545) up to the projection
        coorG(1,:) = X(1:4); coorG(2,:) = Y(1:4); coorG(3,:) = Z(1:4)
        tmp  = 0.5_dp*(coorG(:,3) + coorG(:,4))
        trel = trans3P(coorG(:,1), coorG(:,2), tmp, 6, ierr)
        Tel  = transpose(trel(1:3,1:3))
        do i = 1, 4
           coorG(:,i) = coorG(:,i) - (/X(1),Y(1),Z(1)/)
        end do
        do i = 1, 4
           Xl(i) = dot_product(Tel(1,:),coorG(:,i))
           Yl(i) = dot_product(Tel(2,:),coorG(:,i))
           Zl(i) = dot_product(Tel(3,:),coorG(:,i))
        end do
        call isoMat2D (E,nu,C3)
        Cmat = 0.0_dp
        do i = 1, 3
           do j = 1, 3
              Cmat(i,j)     = C3(i,j)*t
              Cmat(i+3,j+3) = C3(i,j)*t*t*t/12.0_dp
           end do
        end do
        call Andes4shell_stiffmat (Xl,Yl,Zl,Cmat,1.5_dp,0.9_dp,2,Kloc,6,ierr)
        call pMatStiff (coorG,P,6,ierr)
        write(*,'(I4)') ierr
        write(*,'(ES25.17)') trel
        write(*,'(ES25.17)') Kloc
        write(*,'(ES25.17)') P
     case ('fqs35')
        read(*,*) rho, X(1:4), Y(1:4), Z(1:4), TH
        call FQS35 (EMV,X,Y,Z,TH,rho)
        write(*,'(I4)') 0
        write(*,'(ES25.17)') EMV
     case ('quit')
        exit
     case default
        write(*,'(I4)') -99
     end select
  end do
end program ref_driver
