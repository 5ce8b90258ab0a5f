!! Fortran host interface to liblowdin_fock.so.
!!
!! The bind(C) block binds the same symbols, with the same argument passing,
!! as openLOWDIN's own interface block (src/ints/Libint2Interface.f90:165-373):
!! handles by value, arrays as c_loc pointers, the exchange factor by
!! address.  The module procedures follow the reference's Fortran contracts:
!!   Libint2Interface_constructor            (Libint2Interface.f90:382-448)
!!   Libint2Interface_destructor             (Libint2Interface.f90:452-460)
!!   Libint2Interface_compute2BodyIntraspecies_direct (514-547)
!!   Libint2Interface_compute2BodyInterspecies_direct (637-675)
!!   DirectIntegralManager_getDirectIntraRepulsionMatrix
!!   DirectIntegralManager_getDirectInterRepulsionMatrix
!!                                           (DirectIntegralManager.f90:66-179)
!!   Libint2Interface_compute2Body{Intra,Inter}species_direct_IT / _all,
!!   Libint2Interface_computeG12{Intra,Inter}species_direct
!!                                           (Libint2Interface.f90:551-707,
!!                                            935-1179)
!! The reference constructor reads shells from its MolecularSystem; here the
!! shells arrive as flat arrays (angular momentum, primitive count, origin,
!! exponents, raw coefficients), in the same particle-then-contraction order.
module LowdinFockInterface_
  use, intrinsic :: iso_c_binding
  implicit none
  private

  type, public :: Libint2Interface
     type(c_ptr) :: this = c_null_ptr
     logical :: isInstanced = .false.
     integer :: nbasis = 0
  end type Libint2Interface

  interface
     function c_LibintInterface_new(stack_size, id, el, parallel) result(this) bind(C, name="LibintInterface_new")
       import :: c_int, c_bool, c_ptr
       integer(c_int), value :: stack_size
       integer(c_int), value :: id
       logical(c_bool), value :: el
       logical(c_bool), value :: parallel
       type(c_ptr) :: this
     end function c_LibintInterface_new

     subroutine c_LibintInterface_del(this) bind(C, name="LibintInterface_del")
       import :: c_ptr
       type(c_ptr), value :: this
     end subroutine c_LibintInterface_del

     subroutine c_LibintInterface_addShell(this, alpha, coeff, origin, l, nprim) bind(C, name="LibintInterface_add_shell")
       import :: c_ptr, c_int
       type(c_ptr), value :: this
       type(c_ptr), value :: alpha
       type(c_ptr), value :: coeff
       type(c_ptr), value :: origin
       integer(c_int), value :: l
       integer(c_int), value :: nprim
     end subroutine c_LibintInterface_addShell

     function c_LibintInterface_getNbasis(this) result(n) bind(C, name="LibintInterface_get_nbasis")
       import :: c_ptr, c_int
       type(c_ptr), value :: this
       integer(c_int) :: n
     end function c_LibintInterface_getNbasis

     subroutine c_LibintInterface_init2BodyInts(this) bind(C, name="LibintInterface_init_2body_ints")
       import :: c_ptr
       type(c_ptr), value :: this
     end subroutine c_LibintInterface_init2BodyInts

     subroutine c_LibintInterface_compute2BodyDirect(this, density, result, factor) &
          bind(C, name="LibintInterface_compute_2body_direct")
       import :: c_ptr, c_double
       type(c_ptr), value :: this
       type(c_ptr), value :: density
       type(c_ptr), value :: result
       real(c_double) :: factor
     end subroutine c_LibintInterface_compute2BodyDirect

     subroutine c_LibintInterface_computeCouplingDirect(this, othis, density, result) &
          bind(C, name="LibintInterface_compute_coupling_direct")
       import :: c_ptr
       type(c_ptr), value :: this
       type(c_ptr), value :: othis
       type(c_ptr), value :: density
       type(c_ptr), value :: result
     end subroutine c_LibintInterface_computeCouplingDirect

     subroutine c_LibintInterface_compute2BodyDirectIT(this, density, coefficients, result, p) &
          bind(C, name="LibintInterface_compute_2body_directIT")
       import :: c_ptr, c_int
       type(c_ptr), value :: this, density, coefficients, result
       integer(c_int) :: p
     end subroutine c_LibintInterface_compute2BodyDirectIT

     subroutine c_LibintInterface_computeCouplingDirectIT(this, othis, density, coefficients, result, p) &
          bind(C, name="LibintInterface_compute_coupling_directIT")
       import :: c_ptr, c_int
       type(c_ptr), value :: this, othis, density, coefficients, result
       integer(c_int) :: p
     end subroutine c_LibintInterface_computeCouplingDirectIT

     subroutine c_LibintInterface_compute2BodyDirectAll(this, density, result) &
          bind(C, name="LibintInterface_compute_2body_directAll")
       import :: c_ptr
       type(c_ptr), value :: this, density, result
     end subroutine c_LibintInterface_compute2BodyDirectAll

     subroutine c_LibintInterface_computeCouplingDirectAll(this, othis, density, result) &
          bind(C, name="LibintInterface_compute_coupling_directAll")
       import :: c_ptr
       type(c_ptr), value :: this, othis, density, result
     end subroutine c_LibintInterface_computeCouplingDirectAll

     subroutine c_LibintInterface_computeG12Direct(this, density, result, factor, coefficients, exponents, pot_size) &
          bind(C, name="libintinterface_compute_g12_direct")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: this, density, result, coefficients, exponents
       integer(c_int), value :: pot_size
       real(c_double) :: factor
     end subroutine c_LibintInterface_computeG12Direct

     subroutine c_LibintInterface_computeG12InterDirect(this, othis, density, result, coefficients, exponents, &
          pot_size) bind(C, name="libintinterface_compute_g12inter_direct")
       import :: c_ptr, c_int
       type(c_ptr), value :: this, othis, density, result, coefficients, exponents
       integer(c_int), value :: pot_size
     end subroutine c_LibintInterface_computeG12InterDirect

     function c_lowdin_fock_last_error() result(msg) bind(C, name="lowdin_fock_last_error")
       import :: c_ptr
       type(c_ptr) :: msg
     end function c_lowdin_fock_last_error

     function c_strlen(s) result(n) bind(C, name="strlen")
       import :: c_ptr, c_size_t
       type(c_ptr), value :: s
       integer(c_size_t) :: n
     end function c_strlen
  end interface

  public :: Libint2Interface_constructor, Libint2Interface_destructor, &
       Libint2Interface_compute2BodyIntraspecies_direct, Libint2Interface_compute2BodyInterspecies_direct, &
       DirectIntegralManager_getDirectIntraRepulsionMatrix, DirectIntegralManager_getDirectInterRepulsionMatrix, &
       Libint2Interface_compute2BodyIntraspecies_direct_IT, Libint2Interface_compute2BodyInterspecies_direct_IT, &
       Libint2Interface_compute2BodyIntraspecies_direct_all, Libint2Interface_compute2BodyInterspecies_direct_all, &
       Libint2Interface_computeG12Intraspecies_direct, Libint2Interface_computeG12Interspecies_direct, &
       LowdinFock_checkError

contains

  !> Stop with the library's message when its last call failed (the
  !! reference's failures STOP too, DirectIntegralManager.f90:112).
  subroutine LowdinFock_checkError(where)
    character(*), intent(in) :: where
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, k
    character(len=512) :: msg
    p = c_lowdin_fock_last_error()
    n = int(c_strlen(p))
    if (n == 0) return
    call c_f_pointer(p, chars, [n])
    msg = ""
    do k = 1, min(n, 512)
       msg(k:k) = chars(k)
    end do
    write(*, '(A)') "lowdin_fock error in "//trim(where)//": "//trim(msg)
    stop 1
  end subroutine LowdinFock_checkError

  !> Stops (as the reference's callers stop on errors) when an array the C
  !! side reads or writes in full holds fewer than `need` elements.
  subroutine LowdinFock_checkSize(where, what, have, need)
    character(*), intent(in) :: where, what
    integer(8), intent(in) :: have, need
    if (have >= need) return
    write(*, '(A,I0,A,I0)') "lowdin_fock error in "//trim(where)//": "//trim(what)//" holds ", have, &
         " elements, needs ", need
    stop 1
  end subroutine LowdinFock_checkSize

  !> One handle per species, shells in particle then contraction order with
  !! raw coefficients (Libint2Interface.f90:382-448).
  subroutine Libint2Interface_constructor(this, speciesID, isElectron, angularMoment, nprim, origins, &
       exponents, coefficients, parallel)
    type(Libint2Interface), intent(inout) :: this
    integer, intent(in) :: speciesID
    logical, intent(in) :: isElectron
    integer, intent(in) :: angularMoment(:), nprim(:)
    real(8), intent(in) :: origins(:,:)            !! (3, nshell), bohr
    real(8), intent(in) :: exponents(:), coefficients(:)
    logical, optional, intent(in) :: parallel
    real(8), allocatable, target :: e(:), c(:)
    real(8), target :: o(3)
    logical(c_bool) :: par
    integer :: s, off

    par = .true.
    if (present(parallel)) par = parallel
    this%this = c_LibintInterface_new(30000_c_int, int(speciesID, c_int), logical(isElectron, c_bool), par)
    off = 0
    do s = 1, size(angularMoment)
       allocate(e(nprim(s)), c(nprim(s)))
       e = exponents(off + 1:off + nprim(s))
       c = coefficients(off + 1:off + nprim(s))
       o = origins(:, s)
       call c_LibintInterface_addShell(this%this, c_loc(e(1)), c_loc(c(1)), c_loc(o(1)), &
            int(angularMoment(s), c_int), int(nprim(s), c_int))
       call LowdinFock_checkError("add_shell")
       deallocate(e, c)
       off = off + nprim(s)
    end do
    this%nbasis = int(c_LibintInterface_getNbasis(this%this))
    this%isInstanced = .true.
  end subroutine Libint2Interface_constructor

  subroutine Libint2Interface_destructor(this)
    type(Libint2Interface), intent(inout) :: this
    if (this%isInstanced) call c_LibintInterface_del(this%this)
    this%this = c_null_ptr
    this%isInstanced = .false.
  end subroutine Libint2Interface_destructor

  !> Allocates twoBody, rebuilds the pair list, computes G = J + factor K
  !! (Libint2Interface.f90:514-547).
  subroutine Libint2Interface_compute2BodyIntraspecies_direct(density, twoBody, factor, this)
    real(8), allocatable, target, intent(inout) :: density(:,:)
    real(8), allocatable, target, intent(inout) :: twoBody(:,:)
    real(8), intent(inout) :: factor
    type(Libint2Interface), intent(inout) :: this
    real(c_double), target :: f
    if (allocated(twoBody)) deallocate(twoBody)
    allocate(twoBody(this%nbasis, this%nbasis))
    f = factor
    call c_LibintInterface_init2BodyInts(this%this)
    call LowdinFock_checkError("init_2body_ints")
    call c_LibintInterface_compute2BodyDirect(this%this, c_loc(density(1,1)), c_loc(twoBody(1,1)), f)
    call LowdinFock_checkError("compute_2body_direct")
  end subroutine Libint2Interface_compute2BodyIntraspecies_direct

  !> Allocates coupling (n_this x n_this) and computes the Coulomb matrix on
  !! `this` from the other species' density (Libint2Interface.f90:637-675).
  subroutine Libint2Interface_compute2BodyInterspecies_direct(density, coupling, this, other)
    real(8), allocatable, target, intent(inout) :: density(:,:)
    real(8), allocatable, target, intent(inout) :: coupling(:,:)
    type(Libint2Interface), intent(inout) :: this, other
    if (allocated(coupling)) deallocate(coupling)
    allocate(coupling(this%nbasis, this%nbasis))
    call c_LibintInterface_computeCouplingDirect(this%this, other%this, c_loc(density(1,1)), c_loc(coupling(1,1)))
    call LowdinFock_checkError("compute_coupling_direct")
  end subroutine Libint2Interface_compute2BodyInterspecies_direct

  !> Allocates matrixA (n, n, n) and computes the first-quarter transform on
  !! coefficient row p: matrixA(l, k, j) = sum_i C(p, i) (i j|k l), the C
  !! image of result[j n^2 + k n + l] (Libint2Interface.f90:551-579).
  subroutine Libint2Interface_compute2BodyIntraspecies_direct_IT(density, coefficients, matrixA, p, this)
    real(8), target, intent(in) :: density(:,:), coefficients(:,:)
    real(8), allocatable, target, intent(inout) :: matrixA(:,:,:)
    integer, intent(in) :: p
    type(Libint2Interface), intent(inout) :: this
    integer(c_int) :: pp
    call LowdinFock_checkSize("compute_2body_directIT", "density", size(density, kind=8), int(this%nbasis, 8)**2)
    call LowdinFock_checkSize("compute_2body_directIT", "coefficients", size(coefficients, kind=8), &
         int(this%nbasis, 8)**2)
    if (allocated(matrixA)) deallocate(matrixA)
    allocate(matrixA(this%nbasis, this%nbasis, this%nbasis))
    matrixA = 0.0_8
    pp = int(p, c_int)
    call c_LibintInterface_compute2BodyDirectIT(this%this, c_loc(density(1,1)), c_loc(coefficients(1,1)), &
         c_loc(matrixA(1,1,1)), pp)
    call LowdinFock_checkError("compute_2body_directIT")
  end subroutine Libint2Interface_compute2BodyIntraspecies_direct_IT

  !> coupling (n_other, n_other, n_this) (Libint2Interface.f90:677-707).
  subroutine Libint2Interface_compute2BodyInterspecies_direct_IT(density, coefficients, coupling, p, this, other)
    real(8), target, intent(in) :: density(:,:), coefficients(:,:)
    real(8), allocatable, target, intent(inout) :: coupling(:,:,:)
    integer, intent(in) :: p
    type(Libint2Interface), intent(inout) :: this, other
    integer(c_int) :: pp
    call LowdinFock_checkSize("compute_coupling_directIT", "density", size(density, kind=8), int(this%nbasis, 8)**2)
    call LowdinFock_checkSize("compute_coupling_directIT", "coefficients", size(coefficients, kind=8), &
         int(this%nbasis, 8)**2)
    if (allocated(coupling)) deallocate(coupling)
    allocate(coupling(other%nbasis, other%nbasis, this%nbasis))
    pp = int(p, c_int)
    call c_LibintInterface_computeCouplingDirectIT(this%this, other%this, c_loc(density(1,1)), &
         c_loc(coefficients(1,1)), c_loc(coupling(1,1,1)), pp)
    call LowdinFock_checkError("compute_coupling_directIT")
  end subroutine Libint2Interface_compute2BodyInterspecies_direct_IT

  !> ints (n, n, n, n), caller-allocated: the Coulomb branch of
  !! Libint2Interface.f90:1065-1115.
  subroutine Libint2Interface_compute2BodyIntraspecies_direct_all(density, ints, this)
    real(8), target, intent(in) :: density(:,:)
    real(8), target, intent(inout) :: ints(:,:,:,:)
    type(Libint2Interface), intent(inout) :: this
    call LowdinFock_checkSize("compute_2body_directAll", "density", size(density, kind=8), int(this%nbasis, 8)**2)
    call LowdinFock_checkSize("compute_2body_directAll", "ints", size(ints, kind=8), int(this%nbasis, 8)**4)
    call c_LibintInterface_compute2BodyDirectAll(this%this, c_loc(density(1,1)), c_loc(ints(1,1,1,1)))
    call LowdinFock_checkError("compute_2body_directAll")
  end subroutine Libint2Interface_compute2BodyIntraspecies_direct_all

  !> The Coulomb branch of Libint2Interface.f90:1117-1179.
  subroutine Libint2Interface_compute2BodyInterspecies_direct_all(density, ints, this, other)
    real(8), target, intent(in) :: density(:,:)
    real(8), target, intent(inout) :: ints(:,:,:,:)
    type(Libint2Interface), intent(inout) :: this, other
    call LowdinFock_checkSize("compute_coupling_directAll", "ints", size(ints, kind=8), &
         int(this%nbasis, 8)**2 * int(other%nbasis, 8)**2)
    call c_LibintInterface_computeCouplingDirectAll(this%this, other%this, c_loc(density(1,1)), c_loc(ints(1,1,1,1)))
    call LowdinFock_checkError("compute_coupling_directAll")
  end subroutine Libint2Interface_compute2BodyInterspecies_direct_all

  !> Allocates twoBody, rebuilds the pair list, computes J + factor K with
  !! the geminal sum_k coefficients(k) exp(-exponents(k) r12^2)
  !! (Libint2Interface.f90:935-992; the potential's components are passed
  !! in instead of looked up in the InterPotential table).
  subroutine Libint2Interface_computeG12Intraspecies_direct(density, twoBody, factor, coefficients, exponents, this)
    real(8), target, intent(in) :: density(:,:)
    real(8), allocatable, target, intent(inout) :: twoBody(:,:)
    real(8), intent(in) :: factor
    real(8), target, intent(in) :: coefficients(:), exponents(:)
    type(Libint2Interface), intent(inout) :: this
    real(c_double) :: f
    if (allocated(twoBody)) deallocate(twoBody)
    allocate(twoBody(this%nbasis, this%nbasis))
    f = factor
    call c_LibintInterface_init2BodyInts(this%this)
    call LowdinFock_checkError("init_2body_ints")
    call c_LibintInterface_computeG12Direct(this%this, c_loc(density(1,1)), c_loc(twoBody(1,1)), f, &
         c_loc(coefficients(1)), c_loc(exponents(1)), int(size(coefficients), c_int))
    call LowdinFock_checkError("compute_g12_direct")
  end subroutine Libint2Interface_computeG12Intraspecies_direct

  !> Allocates coupling (n_this x n_this): the geminal coupling on `this`
  !! from the other species' density (Libint2Interface.f90:995-1061).
  subroutine Libint2Interface_computeG12Interspecies_direct(density, coupling, coefficients, exponents, this, other)
    real(8), target, intent(in) :: density(:,:)
    real(8), allocatable, target, intent(inout) :: coupling(:,:)
    real(8), target, intent(in) :: coefficients(:), exponents(:)
    type(Libint2Interface), intent(inout) :: this, other
    if (allocated(coupling)) deallocate(coupling)
    allocate(coupling(this%nbasis, this%nbasis))
    call c_LibintInterface_computeG12InterDirect(this%this, other%this, c_loc(density(1,1)), c_loc(coupling(1,1)), &
         c_loc(coefficients(1)), c_loc(exponents(1)), int(size(coefficients), c_int))
    call LowdinFock_checkError("compute_g12inter_direct")
  end subroutine Libint2Interface_computeG12Interspecies_direct

  !> DirectIntegralManager.f90:66-130: copy D, require the LIBINT scheme,
  !! call the species' handle.
  subroutine DirectIntegralManager_getDirectIntraRepulsionMatrix(speciesID, scheme, densityMatrix, &
       twoParticlesMatrix, factor, Libint2Local)
    integer, intent(in) :: speciesID
    character(*), intent(in) :: scheme
    real(8), intent(in) :: densityMatrix(:,:)
    real(8), allocatable, target, intent(inout) :: twoParticlesMatrix(:,:)
    real(8), intent(inout) :: factor
    type(Libint2Interface), intent(inout) :: Libint2Local(:)
    real(8), allocatable, target :: density(:,:)
    if (trim(scheme) /= "LIBINT") stop "The integral method selected has not been implemented"
    allocate(density(size(densityMatrix, 1), size(densityMatrix, 2)))
    density = densityMatrix
    call Libint2Interface_compute2BodyIntraspecies_direct(density, twoParticlesMatrix, factor, &
         Libint2Local(speciesID))
    deallocate(density)
  end subroutine DirectIntegralManager_getDirectIntraRepulsionMatrix

  !> DirectIntegralManager.f90:139-179.
  subroutine DirectIntegralManager_getDirectInterRepulsionMatrix(speciesID, otherSpeciesID, scheme, &
       densityMatrix, couplingMatrix, Libint2Local)
    integer, intent(in) :: speciesID, otherSpeciesID
    character(*), intent(in) :: scheme
    real(8), intent(in) :: densityMatrix(:,:)
    real(8), allocatable, target, intent(inout) :: couplingMatrix(:,:)
    type(Libint2Interface), intent(inout) :: Libint2Local(:)
    real(8), allocatable, target :: density(:,:)
    if (trim(scheme) /= "LIBINT") stop "The integral method selected has not been implemented"
    allocate(density(size(densityMatrix, 1), size(densityMatrix, 2)))
    density = densityMatrix
    call Libint2Interface_compute2BodyInterspecies_direct(density, couplingMatrix, &
         Libint2Local(speciesID), Libint2Local(otherSpeciesID))
    deallocate(density)
  end subroutine DirectIntegralManager_getDirectInterRepulsionMatrix

end module LowdinFockInterface_
