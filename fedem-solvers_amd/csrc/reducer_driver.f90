!> Fortran host driver of the device part reducer.
!>
!> fgpu_reducer_run follows the phase sequence of the reference reducer
!> (src/vpmReducer/reducer.f90:98-716): options, FE part input, SAM and
!> equation numbering, assembly, factorisation with the Schur complement,
!> component modes (EIGVAL), CMSTRS and GRAV, then the _S/_M/_V/_B/_E/_G .fmx
!> and _SAM.fsm files and the .res report.  All arithmetic runs on the GPU
!> behind the C ABI of include/fedem_reducer_gpu.h; this driver only moves
!> options, sizes and result arrays.
!>
!> fgpu_solver_run is the direct static solve of the part under gravity
!> (src/vpmReducer/solver.f90:30-383): the external nodes are clamped, every
!> free DOF is an internal equation, and K u = M a is solved on the resident
!> factor.

module fgpu_bind
  use iso_c_binding
  implicit none

  interface

     integer(c_int) function fgpu_opt_int (name) bind(C,name='fgpu_opt_int')
       import :: c_int, c_char
       character(kind=c_char), intent(in) :: name(*)
     end function fgpu_opt_int

     real(c_double) function fgpu_opt_double (name) bind(C,name='fgpu_opt_double')
       import :: c_double, c_char
       character(kind=c_char), intent(in) :: name(*)
     end function fgpu_opt_double

     integer(c_int) function fgpu_opt_bool (name) bind(C,name='fgpu_opt_bool')
       import :: c_int, c_char
       character(kind=c_char), intent(in) :: name(*)
     end function fgpu_opt_bool

     integer(c_int) function fgpu_opt_isset (name) bind(C,name='fgpu_opt_isset')
       import :: c_int, c_char
       character(kind=c_char), intent(in) :: name(*)
     end function fgpu_opt_isset

     integer(c_int) function fgpu_opt_string (name,buf,len) bind(C,name='fgpu_opt_string')
       import :: c_int, c_char
       character(kind=c_char), intent(in)  :: name(*)
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_int), value :: len
     end function fgpu_opt_string

     integer(c_int) function fgpu_opt_doubles (name,v,n) bind(C,name='fgpu_opt_doubles')
       import :: c_int, c_char, c_double
       character(kind=c_char), intent(in) :: name(*)
       real(c_double), intent(out) :: v(*)
       integer(c_int), value :: n
     end function fgpu_opt_doubles

     integer(c_int) function fgpu_link_read (path) bind(C,name='fgpu_link_read')
       import :: c_int, c_char
       character(kind=c_char), intent(in) :: path(*)
     end function fgpu_link_read

     integer(c_int) function fgpu_link_error (buf,len) bind(C,name='fgpu_link_error')
       import :: c_int, c_char
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_int), value :: len
     end function fgpu_link_error

     integer(c_int) function fgpu_link_set_external (id) bind(C,name='fgpu_link_set_external')
       import :: c_int
       integer(c_int), value :: id
     end function fgpu_link_set_external

     integer(c_int) function fgpu_link_clamp_external () bind(C,name='fgpu_link_clamp_external')
       import :: c_int
     end function fgpu_link_clamp_external

     integer(c_int) function fgpu_link_sizes (sizes) bind(C,name='fgpu_link_sizes')
       import :: c_int
       integer(c_int), intent(out) :: sizes(*)
     end function fgpu_link_sizes

     integer(c_int) function fgpu_link_node_ids (ids) bind(C,name='fgpu_link_node_ids')
       import :: c_int
       integer(c_int), intent(out) :: ids(*)
     end function fgpu_link_node_ids

     integer(c_int) function fgpu_link_attach (h) bind(C,name='fgpu_link_attach')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function fgpu_link_attach

     integer(c_int) function fgpu_link_checksum (ngen,neval,tole,tolf) bind(C,name='fgpu_link_checksum')
       import :: c_int, c_double
       integer(c_int), value :: ngen, neval
       real(c_double), value :: tole, tolf
     end function fgpu_link_checksum

     integer(c_int) function fgpu_write_fmx (file,dtype,cs,data,n) bind(C,name='fgpu_write_fmx')
       import :: c_int, c_char, c_ptr, c_int64_t
       character(kind=c_char), intent(in) :: file(*), dtype(*)
       integer(c_int), value :: cs
       type(c_ptr), value :: data
       integer(c_int64_t), value :: n
     end function fgpu_write_fmx

     integer(c_int) function fgpu_write_sam (h,file,cs,linkid,neval,ngen) bind(C,name='fgpu_write_sam')
       import :: c_int, c_char, c_ptr
       type(c_ptr), value :: h
       character(kind=c_char), intent(in) :: file(*)
       integer(c_int), value :: cs, linkid, neval, ngen
     end function fgpu_write_sam

     integer(c_int) function fgpu_copy_error (h,buf,len) bind(C,name='fgpu_copy_error')
       import :: c_int, c_char, c_ptr
       type(c_ptr), value :: h
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_int), value :: len
     end function fgpu_copy_error

     integer(c_int) function fgpu_create (h) bind(C,name='fgpu_create')
       import :: c_int, c_ptr
       type(c_ptr), intent(out) :: h
     end function fgpu_create

     subroutine fgpu_destroy (h) bind(C,name='fgpu_destroy')
       import :: c_ptr
       type(c_ptr), value :: h
     end subroutine fgpu_destroy

     integer(c_int) function fgpu_set_option (h,name,v) bind(C,name='fgpu_set_option')
       import :: c_int, c_ptr, c_char, c_double
       type(c_ptr), value :: h
       character(kind=c_char), intent(in) :: name(*)
       real(c_double), value :: v
     end function fgpu_set_option

     integer(c_int) function fgpu_analyze (h,ngen,neval,flags) bind(C,name='fgpu_analyze')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: ngen, neval, flags
     end function fgpu_analyze

     integer(c_int) function fgpu_get_sizes (h,sizes) bind(C,name='fgpu_get_sizes')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), intent(out) :: sizes(*)
     end function fgpu_get_sizes

     integer(c_int) function fgpu_get_sam (h,madof,msc,meqn,meqn1,meqn2,pos2) bind(C,name='fgpu_get_sam')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, madof, msc, meqn, meqn1, meqn2, pos2
     end function fgpu_get_sam

     integer(c_int) function fgpu_reduce (h) bind(C,name='fgpu_reduce')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function fgpu_reduce

     integer(c_int) function fgpu_get_reduced (h,sk,sm,g) bind(C,name='fgpu_get_reduced')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, sk, sm, g
     end function fgpu_get_reduced

     integer(c_int) function fgpu_get_eigen (h,nev,lambda,evl) bind(C,name='fgpu_get_eigen')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), intent(out) :: nev
       type(c_ptr), value :: lambda, evl
     end function fgpu_get_eigen

     integer(c_int) function fgpu_get_B (h,B) bind(C,name='fgpu_get_B')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, B
     end function fgpu_get_B

     integer(c_int) function fgpu_get_phi (h,phi) bind(C,name='fgpu_get_phi')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, phi
     end function fgpu_get_phi

     integer(c_int) function fgpu_get_vgi (h,vgi) bind(C,name='fgpu_get_vgi')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, vgi
     end function fgpu_get_vgi

     integer(c_int) function fgpu_get_fe_mass (h,smass) bind(C,name='fgpu_get_fe_mass')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: smass
     end function fgpu_get_fe_mass

     integer(c_int) function fgpu_get_ext_nodes (h,nenod,xyz,medof) bind(C,name='fgpu_get_ext_nodes')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, xyz, medof
       integer(c_int), intent(out) :: nenod
     end function fgpu_get_ext_nodes

     integer(c_int) function fgpu_reduced_eigen (n,sk,sm,nval,lambda) bind(C,name='fgpu_reduced_eigen')
       import :: c_int, c_double
       integer(c_int), value :: n, nval
       real(c_double), intent(in)  :: sk(*), sm(*)
       real(c_double), intent(out) :: lambda(*)
     end function fgpu_reduced_eigen

     integer(c_int) function fgpu_get_gravity_internal (h,gi) bind(C,name='fgpu_get_gravity_internal')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, gi
     end function fgpu_get_gravity_internal

     integer(c_int) function fgpu_reduce_loads (h,nrhs,F,Fred) bind(C,name='fgpu_reduce_loads')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: nrhs
       type(c_ptr), value :: F, Fred
     end function fgpu_reduce_loads

     integer(c_int) function fgpu_link_loads (lcase,node,dof,val) bind(C,name='fgpu_link_loads')
       import :: c_int, c_ptr
       type(c_ptr), value :: lcase, node, dof, val
     end function fgpu_link_loads

     integer(c_int) function fgpu_get_pivots (h,counts,node,ldof,eqn,kind,nmax) bind(C,name='fgpu_get_pivots')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, node, ldof, eqn, kind
       integer(c_int), intent(out) :: counts(3)
       integer(c_int), value :: nmax
     end function fgpu_get_pivots

     integer(c_int) function fgpu_solve_internal (h,nrhs,R,X) bind(C,name='fgpu_solve_internal')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: nrhs
       type(c_ptr), value :: R, X
     end function fgpu_solve_internal

  end interface

  integer, parameter :: dp = c_double
  real(dp), parameter :: pi_p = 3.14159265358979323846_dp

contains

  !> @brief C string of a Fortran string.
  function cs (s) result(c)
    character(len=*), intent(in) :: s
    character(len=len_trim(s)+1,kind=c_char) :: c
    c = trim(s)//c_null_char
  end function cs

  !> @brief String value of a command-line option.
  subroutine opt_string (name,val)
    character(len=*), intent(in)  :: name
    character(len=*), intent(out) :: val
    integer :: n
    n = fgpu_opt_string(cs(name),val,len(val))
    if (n < 0) val = ''
  end subroutine opt_string

  !> @brief File name of an output: the option value, else the link file name
  !> with its extension replaced by the suffix (inputReducerModule.f90:37-69).
  subroutine file_name (option,suffix,name)
    character(len=*), intent(in)  :: option, suffix
    character(len=*), intent(out) :: name
    integer :: idot
    call opt_string (option,name)
    if (name /= '' .or. suffix == '') return
    call opt_string ('linkfile',name)
    if (name == '') then
       name = 'reducer'//suffix
       return
    end if
    idot = index(name,'.',.true.)
    if (idot < 1) idot = len_trim(name)+1
    name(idot:) = suffix
  end subroutine file_name

  !> @brief Reads the link file and the extra external nodes of -extNodes.
  subroutine read_link (lpu,ierr)
    integer, intent(in)  :: lpu
    integer, intent(out) :: ierr
    character(len=4096) :: path, ext, msg
    integer :: i, j, id, rc
    call opt_string ('linkfile',path)
    if (path == '') then
       write(lpu,"(' *** Error: No link file (-linkfile) given')")
       ierr = -1
       return
    end if
    ierr = fgpu_link_read(cs(path))
    if (ierr /= 0) then
       rc = fgpu_link_error(msg,len(msg))
       write(lpu,"(' *** Error: ',A)") trim(msg)
       return
    end if
    ! -extNodes: node ids separated by blanks or commas, brackets ignored
    call opt_string ('extNodes',ext)
    do i = 1, len_trim(ext)
       if (index('0123456789',ext(i:i)) < 1) ext(i:i) = ' '
    end do
    i = 1
    do while (i <= len_trim(ext))
       if (ext(i:i) == ' ') then
          i = i + 1
          cycle
       end if
       j = i
       do while (j < len_trim(ext) .and. ext(j+1:j+1) /= ' ')
          j = j + 1
       end do
       read(ext(i:j),*) id
       ierr = fgpu_link_set_external(int(id,c_int))
       if (ierr /= 0) then
          rc = fgpu_link_error(msg,len(msg))
          write(lpu,"(' *** Error: ',A)") trim(msg)
          return
       end if
       i = j + 1
    end do
  end subroutine read_link

  !> @brief Assembles the nodal loads of the link into full load vectors
  !> F(neq,nlc) in equation order (rhsINADD, inadd.f90:467-604, for nodal
  !> point loads): loads on fixed or dependent DOFs are dropped.
  subroutine build_loads (h,nnod,ndof,neq,nlc,F)
    type(c_ptr), intent(in) :: h
    integer, intent(in)     :: nnod, ndof, neq, nlc
    real(dp), intent(out)   :: F(neq,nlc)
    integer, allocatable, target  :: madof(:), meqn(:), lc(:), nod(:), idof(:)
    real(dp), allocatable, target :: val(:)
    integer :: nl, k, ieq, rc
    F = 0.0_dp
    nl = fgpu_link_loads(c_null_ptr,c_null_ptr,c_null_ptr,c_null_ptr)
    if (nl <= 0 .or. nlc <= 0) return
    allocate(madof(nnod+1),meqn(ndof),lc(nl),nod(nl),idof(nl),val(nl))
    rc = fgpu_get_sam(h,c_loc(madof),c_null_ptr,c_loc(meqn),c_null_ptr,c_null_ptr,c_null_ptr)
    nl = fgpu_link_loads(c_loc(lc),c_loc(nod),c_loc(idof),c_loc(val))
    do k = 1, nl
       if (madof(nod(k))+idof(k)-1 >= madof(nod(k)+1)) cycle
       ieq = meqn(madof(nod(k))+idof(k)-1)
       if (ieq > 0 .and. ieq <= neq .and. lc(k) <= nlc) F(ieq,lc(k)) = F(ieq,lc(k)) + val(k)
    end do
  end subroutine build_loads

  !> @brief Reports a failure to hand the FE part to the device.
  subroutine link_failure (lpu,ierr)
    integer, intent(in) :: lpu, ierr
    character(len=1024) :: msg
    integer :: rc
    rc = fgpu_link_error(msg,len(msg))
    write(lpu,"(' *** Error: Part input failed, ierr =',I4/5X,A)") ierr, trim(msg)
  end subroutine link_failure

  !> @brief Reports the last device error of a handle.
  subroutine report (lpu,h,what,ierr)
    integer, intent(in)          :: lpu, ierr
    type(c_ptr), intent(in)      :: h
    character(len=*), intent(in) :: what
    character(len=1024) :: msg
    integer :: rc
    rc = fgpu_copy_error(h,msg,len(msg))
    write(lpu,"(' *** Error: ',A,' failed, ierr =',I4/5X,A)") what, ierr, trim(msg)
  end subroutine report

  !> @brief Reports the singular and negative pivots of the last
  !> factorisation (reportSingEqns, cmstrs.f90:1324-1352; CMSTRS,
  !> cmstrs.f90:644-652; warningOnNegativePivots,
  !> solExtensionModule.f90:1023-1031).  Node numbers are the part's ids.
  subroutine pivot_report (lpu,h,ierr)
    integer, intent(in)     :: lpu, ierr
    type(c_ptr), intent(in) :: h
    integer(c_int) :: counts(3), nsing, nnod, rc
    integer(c_int), allocatable, target :: node(:), ldof(:), eqn(:), kind(:), ids(:)
    integer(c_int) :: szs(4)
    integer :: i
    counts = 0
    nsing = fgpu_get_pivots(h,counts,c_null_ptr,c_null_ptr,c_null_ptr,c_null_ptr,0_c_int)
    if (nsing > 0) then
       allocate(node(nsing),ldof(nsing),eqn(nsing),kind(nsing))
       nsing = fgpu_get_pivots(h,counts,c_loc(node),c_loc(ldof),c_loc(eqn),c_loc(kind),nsing)
       rc = fgpu_link_sizes(szs)
       nnod = szs(1)
       allocate(ids(max(nnod,1)))
       rc = fgpu_link_node_ids(ids)
       write(lpu,"(/' *** Error: Singular or ill-conditioned stiffness matrix detected at the following node(s):')")
       do i = 1, nsing
          if (node(i) >= 1 .and. node(i) <= nnod) node(i) = ids(node(i))
          if (kind(i) == 1) then
             write(lpu,"(5X,'Node',i9,' local DOF',i3,' ( eqn. # ',i9,' )  zero pivot')") node(i), ldof(i), eqn(i)
          else
             write(lpu,"(5X,'Node',i9,' local DOF',i3,' ( eqn. # ',i9,' )  reduced-to-zero pivot')") &
                  &   node(i), ldof(i), eqn(i)
          end if
       end do
       if (ierr == 0) write(lpu,"(' *** Warning: The above singularities have been circumvented by replacing',&
            &  ' the zero (or nearly zero)'/5X,'pivots by the magnitude of the original diagonal entry.')")
       deallocate(node,ldof,eqn,kind,ids)
    end if
    if (counts(1) > 0) then
       write(lpu,"(/' *** Warning: The stiffness matrix is not positive definite.'" // &
            &     "/5X,'There are',i6,' negative pivots.'" // &
            &     "/5X,'This can be caused by too many ill-shaped finite elements.'" // &
            &     "/5X,'Using this FE part may lead to an instable dynamics simulation.')") counts(1)
    end if
  end subroutine pivot_report

  !> @brief Writes one column-major FP64 array with its file tag.
  subroutine write_db (lpu,fname,dtype,chk,a,n,ierr)
    integer, intent(in)              :: lpu, chk
    character(len=*), intent(in)     :: fname, dtype
    integer(c_int64_t), intent(in)   :: n
    real(dp), intent(in), target     :: a(*)
    integer, intent(out)             :: ierr
    ierr = fgpu_write_fmx(cs(fname),cs(dtype),int(chk,c_int),c_loc(a),n)
    if (ierr /= 0) write(lpu,"(' *** Error: Can not write ',A,' to ',A)") dtype, trim(fname)
  end subroutine write_db

  !> @brief Mass, centre of gravity and inertia of the reduced part
  !> (JCMS, cmstrs.f90:1083-1184): with R the rigid-body motions of the
  !> external nodes (unit translations, and unit rotations about a point c),
  !> I = R^T SM R.  The mass is the mean of the three translational masses;
  !> the centre of gravity follows from the translation-rotation couplings
  !> about the origin, and the inertia is reported about both points.
  subroutine mass_props (lpu,sm,nenod,tenc,medof,rmass)
    integer , intent(in)  :: lpu, nenod, medof(:)
    real(dp), intent(in)  :: sm(:,:), tenc(:,:)
    real(dp), intent(out) :: rmass
    real(dp), allocatable :: R(:,:)
    real(dp) :: Io(6,6), Ic(6,6), cg(3)
    allocate(R(size(sm,1),6))
    call rigid_modes ([0.0_dp,0.0_dp,0.0_dp])
    Io = matmul(transpose(R),matmul(sm,R))
    rmass = (Io(1,1)+Io(2,2)+Io(3,3)) / 3.0_dp
    if (rmass > epsilon(1.0_dp)) then
       cg = [Io(2,6), Io(3,4), Io(1,5)] / rmass
    else
       cg = 0.0_dp
    end if
    call rigid_modes (cg)
    Ic = matmul(transpose(R),matmul(sm,R))
    write(lpu,600) rmass, cg, Io(4,4), Ic(4,4), Io(5,5), Ic(5,5), Io(6,6), Ic(6,6), &
         &         Io(4,5), Ic(4,5), Io(5,6), Ic(5,6), Io(6,4), Ic(6,4)
600 format(//4X,'MASS AND INERTIA PROPERTIES OF THE REDUCED SYSTEM' &
         &  /4X,49('-') &
         & //4X,'MASS      = ',1PE14.5 &
         &  /4X,'CG X_COOR = ',1PE14.5 /4X,'CG Y_COOR = ',1PE14.5 /4X,'CG Z_COOR = ',1PE14.5 &
         & //4X,'               About origin      About CG' &
         &  /4X,'IXX       = ',1PE14.5,E18.5 /4X,'IYY       = ',1PE14.5,E18.5 &
         &  /4X,'IZZ       = ',1PE14.5,E18.5 /4X,'IXY       = ',1PE14.5,E18.5 &
         &  /4X,'IYZ       = ',1PE14.5,E18.5 /4X,'IZX       = ',1PE14.5,E18.5)
  contains
    subroutine rigid_modes (c)
      real(dp), intent(in) :: c(3)
      real(dp) :: x(3)
      integer  :: i, j, n
      R = 0.0_dp
      do i = 1, nenod
         j = medof(i) - 1
         n = medof(i+1) - medof(i)
         x = tenc(:,i) - c
         R(j+1,1) = 1.0_dp
         R(j+2,2) = 1.0_dp
         R(j+3,3) = 1.0_dp
         R(j+2,4) = -x(3) ! e_x cross x
         R(j+3,4) =  x(2)
         R(j+1,5) =  x(3) ! e_y cross x
         R(j+3,5) = -x(1)
         R(j+1,6) = -x(2) ! e_z cross x
         R(j+2,6) =  x(1)
         if (n >= 4) R(j+4,4) = 1.0_dp
         if (n >= 5) R(j+5,5) = 1.0_dp
         if (n >= 6) R(j+6,6) = 1.0_dp
      end do
    end subroutine rigid_modes
  end subroutine mass_props

  !> @brief Free-free eigenfrequencies of the reduced part (EIGCMS,
  !> cmstrs.f90:1215-1310): the lowest nevred eigenvalues of
  !> SK z = lambda SM z, printed as f = sqrt(lambda)/(2 pi) (negative
  !> eigenvalues as imaginary frequencies).  A mass matrix that is not
  !> positive definite gives the reference's warning and no frequencies.
  subroutine reduced_frequencies (lpu,sk,sm,nevred)
    integer , intent(in) :: lpu, nevred
    real(dp), intent(in) :: sk(:,:), sm(:,:)
    real(dp), allocatable :: lambda(:)
    integer :: i, n, rc
    n = size(sk,1)
    allocate(lambda(nevred))
    rc = fgpu_reduced_eigen(int(n,c_int),sk,sm,int(nevred,c_int),lambda)
    if (rc > n) then
       write(lpu,"(/' *** Error: No eigenvalues can be computed.'/5X,'The leading minor of order',I6, &
            &      ' of the mass matrix is not positive definite.'/5X, &
            &      'Check that you have a proper mass distribution.')") rc - n
    else if (rc /= 0) then
       write(lpu,"(/' *** Error: Eigenvalues of the reduced system failed, rc =',I6)") rc
    end if
    if (rc /= 0) then
       write(lpu,"('  ** Warning: A dynamics simulation using this superelement may fail.')")
       return
    end if
    write(lpu,"(//4X,'EIGEN FREQUENCIES OF THE REDUCED SYSTEM'/4X,39('-')//4X,'MODE  FREQUENCY [Hz]')")
    do i = 1, nevred
       if (lambda(i) < 0.0_dp) then
          write(lpu,"(I7,1PE17.8,'*i')") i, -sqrt(-lambda(i))*0.5_dp/pi_p
       else
          write(lpu,"(I7,1PE17.8)") i, sqrt(lambda(i))*0.5_dp/pi_p
       end if
    end do
  end subroutine reduced_frequencies

end module fgpu_bind


!> @brief The part reduction (reducer.f90 through the device ABI).
subroutine fgpu_reducer_run (ierr) bind(C,name='fgpu_reducer_run')
  use fgpu_bind
  implicit none
  integer(c_int), intent(out) :: ierr

  integer  :: lpu, ngen, neval, nev, ndim, ndof1, ndof2, flags, chk, i, rc
  integer  :: sizes(8), lsz(4), linkId, nlc
  logical  :: noGrav, noMass, calcGdisp
  real(dp) :: tolEigval, tolFactorize, eigenShift
  type(c_ptr) :: h
  character(len=4096) :: fname
  real(dp), allocatable, target :: sk(:,:), sm(:,:), g(:,:), B(:,:), phi(:,:), vgi(:,:)
  real(dp), allocatable, target :: F(:,:), Fred(:,:)
  real(dp), allocatable, target :: lambda(:), evl(:), tenc(:,:)
  integer , allocatable, target :: medof(:)
  real(dp) :: sMass, rMass
  integer  :: nenod, nevred

  h = c_null_ptr
  call file_name ('resfile','.res',fname)
  open(newunit=lpu,file=trim(fname),status='replace',action='write',iostat=ierr)
  if (ierr /= 0) then
     ierr = -1
     return
  end if
  write(lpu,"(/4X,'FEDEM part reducer, MI355X device path'/)")

  ! --- Options (reducer.f90:115-180; inputReducerModule.f90:159-166)
  ngen  = fgpu_opt_int(cs('ngen'))
  if (ngen > 0) then
     neval = max(fgpu_opt_int(cs('neval')),ngen)
  else
     ngen  = 0
     neval = 0
  end if
  linkId    = fgpu_opt_int(cs('linkId'))
  noGrav    = fgpu_opt_bool(cs('nograv')) /= 0
  noMass    = fgpu_opt_bool(cs('nomass')) /= 0 .or. noGrav
  calcGdisp = fgpu_opt_isset(cs('dispfile')) /= 0
  tolEigval    = fgpu_opt_double(cs('tolEigval'))
  tolFactorize = max(fgpu_opt_double(cs('tolFactorize')),1.0e-20_dp)
  eigenShift   = fgpu_opt_double(cs('eigenshift'))
  if (fgpu_opt_bool(cs('lumpedmass')) /= 0) then
     write(lpu,"(' *** Error: -lumpedmass is not supported by the device reducer')")
     ierr = -1
     goto 900
  end if

  ! --- FE part input
  call read_link (lpu,ierr)
  if (ierr /= 0) goto 900
  chk = fgpu_link_checksum(int(ngen,c_int),int(neval,c_int),tolEigval,tolFactorize)

  ierr = fgpu_create(h)
  if (ierr /= 0) goto 900
  rc = fgpu_set_option(h,cs('eigenshift'),eigenShift)
  ! Ritz residual bound of the Lanczos convergence test: the eigenvalue error
  ! of a shift-invert Ritz pair is of the order of the squared residual, so
  ! -tolEigval is met by the default bound 1e-11 (which the modes' own
  ! accuracy needs); a tighter -tolEigval tightens it
  rc = fgpu_set_option(h,cs('lanczosTol'),min(1.0e-11_dp,tolEigval))
  if (fgpu_opt_int(cs('lanczosBlock')) > 0) &
       rc = fgpu_set_option(h,cs('lanczosBlock'),real(fgpu_opt_int(cs('lanczosBlock')),dp))
  if (fgpu_opt_int(cs('lanczosMaxVectors')) > 0) &
       rc = fgpu_set_option(h,cs('lanczosMaxVectors'),real(fgpu_opt_int(cs('lanczosMaxVectors')),dp))
  if (fgpu_opt_bool(cs('mixedPrecision')) /= 0) rc = fgpu_set_option(h,cs('mixedPrecision'),1.0_dp)
  rc = fgpu_set_option(h,cs('tolFactorize'),max(fgpu_opt_double(cs('tolFactorize')),1.0e-20_dp))
  rc = fgpu_set_option(h,cs('singularityHandler'),real(fgpu_opt_int(cs('singularityHandler')),dp))
  ierr = fgpu_link_attach(h)
  if (ierr /= 0) then
     call link_failure (lpu,ierr)
     goto 900
  end if

  ! --- SAM, equation numbering and symbolic analysis
  flags = 0
  if (calcGdisp) flags = flags + 1
  if (noGrav .and. .not. calcGdisp) flags = flags + 2 ! no mass matrix at all
  ierr = fgpu_analyze(h,int(ngen,c_int),int(neval,c_int),int(flags,c_int))
  if (ierr /= 0) then
     call report (lpu,h,'Analysis',ierr)
     goto 900
  end if
  ierr = fgpu_get_sizes(h,sizes)
  ndof1 = sizes(4)
  ndof2 = sizes(5)
  ndim  = sizes(6)
  write(lpu,600) sizes(8),sizes(1),sizes(7),ngen,sizes(2),sizes(3),ndof1,ndof2,ndim
600 format(/4X,'FE MODEL SUMMARY' / 4X,16('-') &
         & //4X,'Number of elements             =',I8 &
         &  /4X,'Number of nodes (total)        =',I8 &
         &  /4X,'Number of external nodes       =',I8 &
         &  /4X,'Number of generalized modes    =',I8 &
         &  /4X,'Number of DOFs                 =',I8 &
         &  /4X,'Number of equations            =',I8 &
         &  /4X,'Number of internal equations   =',I8 &
         &  /4X,'Number of external equations   =',I8 &
         &  /4X,'Dimension of reduced matrices  =',I8 / )

  call file_name ('samfile','_SAM.fsm',fname)
  ierr = fgpu_write_sam(h,cs(fname),int(chk,c_int),int(linkId,c_int),int(neval,c_int),int(ngen,c_int))
  if (ierr /= 0) then
     write(lpu,"(' *** Error: Can not save SAM-data to file ',A)") trim(fname)
     goto 900
  end if

  ! --- The numeric reduction on the device
  ierr = fgpu_reduce(h)
  call pivot_report (lpu,h,ierr)
  if (ierr /= 0) then
     call report (lpu,h,'Reduction',ierr)
     goto 900
  end if

  ! --- Component modes (reducer.f90:390-425)
  if (neval > 0) then
     allocate(lambda(neval),evl(neval),phi(ndof1,neval))
     ierr = fgpu_get_eigen(h,nev,c_loc(lambda),c_loc(evl))
     write(lpu,6010) (i,evl(i)*0.5_dp/pi_p,i=1,nev)
6010 format(/4X,'EIGENFREQUENCIES OF THE COMPONENT MODES [Hz]' &
          &  /(4X,I4,1PE16.7))
     ierr = fgpu_get_phi(h,c_loc(phi))
     call file_name ('eigfile','_E.fmx',fname)
     call write_db (lpu,fname,'generalized modes',chk,phi,int(size(phi),c_int64_t),ierr)
     if (ierr /= 0) goto 900
     deallocate(phi)
  end if

  ! --- Displacement recovery matrix (reducer.f90:367, diskMatrixModule)
  allocate(B(ndof1,ndof2))
  ierr = fgpu_get_B(h,c_loc(B))
  call file_name ('Bmatfile','_B.fmx',fname)
  call write_db (lpu,fname,'disk matrix',chk,B,int(size(B),c_int64_t),ierr)
  deallocate(B)
  if (ierr /= 0) goto 900

  ! --- Reduced matrices (reducer.f90:571-602, 668-675)
  allocate(sk(ndim,ndim),sm(ndim,ndim),g(ndim,3))
  ierr = fgpu_get_reduced(h,c_loc(sk),c_loc(sm),c_loc(g))
  if (ierr /= 0) then
     call report (lpu,h,'Result transfer',ierr)
     goto 900
  end if
  ! --- Mass properties of the reduced part vs the FE mass (reducer.f90:535-567)
  sMass = 0.0_dp
  rMass = 0.0_dp
  if (.not. noMass) then
     rc = fgpu_get_fe_mass(h,sMass)
     rc = fgpu_get_ext_nodes(h,nenod,c_null_ptr,c_null_ptr)
     allocate(tenc(3,max(nenod,1)),medof(nenod+1))
     rc = fgpu_get_ext_nodes(h,nenod,c_loc(tenc),c_loc(medof))
     write(lpu,"(/4X,'TOTAL MASS OF SUBSTRUCTURE :',1PD14.5)") sMass
     call mass_props (lpu,sm,nenod,tenc,medof,rMass)
     deallocate(tenc,medof)
     if (abs(sMass) <= epsilon(1.0_dp)) then
        write(lpu,"(/'  ** Note: This FE part has no mass')")
     else if (abs(rMass-sMass) > 1.0e-6_dp*sMass) then
        write(lpu,"(/'  ** Warning: The total mass computed from the reduced substructure matrices' &
             &      /14X,'deviates from the mass computed during finite element assembly.' &
             &      /14X,'Mass deviation = ',1PE12.5,'  (',2PF6.2,'%)')") rMass-sMass, abs(rMass-sMass)/sMass
     end if
  end if

  call file_name ('stiffile','_S.fmx',fname)
  call write_db (lpu,fname,'stiffness matrix',chk,sk,int(size(sk),c_int64_t),ierr)
  if (ierr /= 0) goto 900
  if (.not. noMass) then
     call file_name ('massfile','_M.fmx',fname)
     call write_db (lpu,fname,'mass matrix',chk,sm,int(size(sm),c_int64_t),ierr)
     if (ierr /= 0) goto 900
  end if
  if (calcGdisp) then
     allocate(vgi(ndof1,3))
     ierr = fgpu_get_vgi(h,c_loc(vgi))
     call file_name ('dispfile','_V.fmx',fname)
     call write_db (lpu,fname,'displacement matrix',chk,vgi,int(size(vgi),c_int64_t),ierr)
     deallocate(vgi)
     if (ierr /= 0) goto 900
  end if
  ! --- Reduced load vectors (reducer.f90:598-607)
  rc = fgpu_link_sizes(lsz)
  nlc = lsz(4)
  if (nlc > 0) then
     allocate(F(sizes(3),nlc),Fred(ndim,nlc))
     call build_loads (h,sizes(1),sizes(2),sizes(3),nlc,F)
     ierr = fgpu_reduce_loads(h,int(nlc,c_int),c_loc(F),c_loc(Fred))
     if (ierr /= 0) then
        call report (lpu,h,'Load vector reduction',ierr)
        goto 900
     end if
     call file_name ('loadfile','_L.fmx',fname)
     call write_db (lpu,fname,'load vectors',chk,Fred,int(size(Fred),c_int64_t),ierr)
     if (ierr /= 0) goto 900
  end if
  ! --- Free-free frequencies of the reduced part (reducer.f90:609-650)
  nevred = fgpu_opt_int(cs('nevred'))
  if (nevred > 0 .and. sMass > epsilon(1.0_dp) .and. rMass > epsilon(1.0_dp)) then
     call reduced_frequencies (lpu,sk,sm,min(nevred,ndim))
  end if
  if (.not. noGrav) then
     call file_name ('gravfile','_G.fmx',fname)
     call write_db (lpu,fname,'gravity force vectors',chk,g,int(size(g),c_int64_t),ierr)
     if (ierr /= 0) goto 900
  end if
  write(lpu,"(/4X,'Reduced matrices written, dimension',I6)") ndim

900 continue
  if (c_associated(h)) call fgpu_destroy (h)
  if (ierr /= 0) write(lpu,"(/' *** Part reduction failed, ierr =',I6)") ierr
  close(lpu)
  if (ierr > 0) ierr = -ierr
end subroutine fgpu_reducer_run


!> @brief Direct static solve of the part under gravity (solver.f90): returns
!> the translations of the n listed nodes (ids) in displ(3,n).
subroutine fgpu_solver_run (n,nodes,displ,ierr) bind(C,name='fgpu_solver_run')
  use fgpu_bind
  implicit none
  integer(c_int), value       :: n
  integer(c_int), intent(in)  :: nodes(*)
  real(c_double), intent(out) :: displ(3,*)
  integer(c_int), intent(out) :: ierr

  integer  :: lpu, i, j, k, nnod, ndof, ndof1, nd, rc, sizes(8), lsz(4), ieq, nlc, nrhs, flags
  real(dp) :: grav(3), dMax
  integer  :: nMax
  type(c_ptr) :: h
  character(len=4096) :: fname
  integer,  allocatable, target :: madof(:), meqn(:), ids(:)
  real(dp), allocatable, target :: gi(:,:), rhs(:,:), u(:,:)

  h = c_null_ptr
  if (n > 0) displ(:,1:n) = 0.0_dp
  call file_name ('resfile','.res',fname)
  open(newunit=lpu,file=trim(fname),status='replace',action='write',iostat=ierr)
  if (ierr /= 0) then
     ierr = -1
     return
  end if
  write(lpu,"(/4X,'FEDEM part solver, MI355X device path'/)")

  grav = 0.0_dp
  nd = fgpu_opt_doubles(cs('gvec'),grav,3_c_int)

  call read_link (lpu,ierr)
  if (ierr /= 0) goto 900
  rc = fgpu_link_clamp_external()
  rc = fgpu_link_sizes(lsz)
  allocate(ids(lsz(1)))
  rc = fgpu_link_node_ids(ids)

  ierr = fgpu_create(h)
  if (ierr /= 0) goto 900
  if (fgpu_opt_bool(cs('mixedPrecision')) /= 0) rc = fgpu_set_option(h,cs('mixedPrecision'),1.0_dp)
  rc = fgpu_set_option(h,cs('tolFactorize'),max(fgpu_opt_double(cs('tolFactorize')),1.0e-20_dp))
  rc = fgpu_set_option(h,cs('singularityHandler'),real(fgpu_opt_int(cs('singularityHandler')),dp))
  ierr = fgpu_link_attach(h)
  if (ierr /= 0) then
     call link_failure (lpu,ierr)
     goto 900
  end if
  ! nrhs = max(load cases, 1 with gravity) (solver.f90:104-108,131-135)
  nlc  = lsz(4)
  nrhs = nlc
  if (sum(grav*grav) > 1.0e-8_dp) nrhs = max(nrhs,1)
  if (nrhs < 1) then
     write(lpu,"(' *** Error: No loads (FORCE/MOMENT cards or -gvec) to solve for')")
     ierr = -1
     goto 900
  end if
  flags = 2 ! stiffness only
  if (sum(grav*grav) > 1.0e-8_dp) flags = 0
  ierr = fgpu_analyze(h,0_c_int,0_c_int,int(flags,c_int))
  if (ierr /= 0) then
     call report (lpu,h,'Analysis',ierr)
     goto 900
  end if
  rc = fgpu_get_sizes(h,sizes)
  nnod  = sizes(1)
  ndof  = sizes(2)
  ndof1 = sizes(4)
  ierr = fgpu_reduce(h)
  call pivot_report (lpu,h,ierr)
  if (ierr /= 0) then
     call report (lpu,h,'Factorization',ierr)
     goto 900
  end if

  ! --- Right-hand sides: nodal loads per load case, plus M a in every case
  allocate(rhs(ndof1,nrhs),u(ndof1,nrhs),madof(nnod+1),meqn(ndof))
  call build_loads (h,nnod,ndof,ndof1,nrhs,rhs)
  if (flags == 0) then
     allocate(gi(ndof1,3))
     ierr = fgpu_get_gravity_internal(h,c_loc(gi))
     if (ierr /= 0) then
        call report (lpu,h,'Gravity forces',ierr)
        goto 900
     end if
     do j = 1, nrhs
        rhs(:,j) = rhs(:,j) + matmul(gi,grav)
     end do
  end if
  ierr = fgpu_solve_internal(h,int(nrhs,c_int),c_loc(rhs),c_loc(u))
  if (ierr /= 0) then
     call report (lpu,h,'Linear solve',ierr)
     goto 900
  end if
  rc = fgpu_get_sam(h,c_loc(madof),c_null_ptr,c_loc(meqn),c_null_ptr,c_null_ptr,c_null_ptr)

  ! --- Nodal translations (solver.f90:298-321); all equations are internal
  do k = 1, 3
     nMax = 0
     dMax = 0.0_dp
     do i = 1, nnod
        if (madof(i)+k > madof(i+1)) cycle
        ieq = meqn(madof(i)+k-1)
        if (ieq < 1 .or. ieq > ndof1) cycle
        if (nMax == 0 .or. abs(u(ieq,1)) > abs(dMax)) then
           nMax = ids(i)
           dMax = u(ieq,1)
        end if
        do j = 1, n
           if (nodes(j) == ids(i)) displ(k,j) = u(ieq,1)
        end do
     end do
     write(lpu,"('    Max ',A1,'-displacement:',1PE12.5,'  in node',I8)") &
          char(ichar('W')+k), dMax, nMax
  end do

900 continue
  if (c_associated(h)) call fgpu_destroy (h)
  if (ierr /= 0) write(lpu,"(/' *** Part solve failed, ierr =',I6)") ierr
  close(lpu)
  if (ierr > 0) ierr = -ierr
end subroutine fgpu_solver_run
