!! Fortran host driver of the direct Fock build: reads a job description
!! (species, shells, densities), performs what one SCF iteration of
!! openLOWDIN's direct path performs for every species --
!!   WaveFunction_buildTwoParticlesMatrix (src/scf/WaveFunction.f90:817-1159):
!!     factor = kappa/eta, G only if N > 1, G <- G q^2
!!   WaveFunction_buildCouplingMatrix (:1163-1535):
!!     for every other species: coupling <- coupling q_i q_j, upper triangle
!!     copied to the lower one, summed
!! -- through the ISO_C_BINDING layer, and writes the raw library results and
!! the scaled species matrices as raw FP64 column-major files.
!!
!! Usage: fock_host_driver <job-directory>
!!   job.txt:  nspecies
!!             per species: id is_electron charge kappa eta nparticles nshell
!!                          per shell: l nprim x y z / exponents / coefficients
!!   dens_<id>.bin: density of species <id> (n x n, column-major FP64)
!! Writes g_<id>.bin (raw J + factor K), c_<id>_<jd>.bin (raw coupling on id
!! from jd), two_<id>.bin and coupling_<id>.bin (scaled, as the SCF uses them).
program fock_host_driver
  use LowdinFockInterface_
  implicit none
  type(Libint2Interface), allocatable :: objs(:)
  character(len=1024) :: dir
  integer :: ns, k, j, s, u, nsh, np_total, isel
  integer, allocatable :: ids(:), npart(:), nbas(:)
  real(8), allocatable :: charge(:), kappa(:), eta(:)
  type :: mat
     real(8), allocatable :: v(:,:)
  end type mat
  type(mat), allocatable :: dens(:)
  integer, allocatable :: lam(:), npr(:)
  real(8), allocatable :: orig(:,:), ex(:), co(:), tmpe(:), tmpc(:)
  real(8), allocatable, target :: G(:,:), aux(:,:), coupling(:,:)
  real(8) :: factor
  integer :: i, i2, ng, p
  logical :: have
  real(8), allocatable :: gexp(:), gcoef(:)
  real(8), allocatable, target :: A3(:,:,:), A4(:,:,:,:)
  integer :: ghost_id

  call get_command_argument(1, dir)
  open(newunit=u, file=trim(dir)//"/job.txt", status="old", action="read")
  read(u, *) ns
  allocate(objs(ns), ids(ns), npart(ns), nbas(ns), charge(ns), kappa(ns), eta(ns), dens(ns))
  do k = 1, ns
     read(u, *) ids(k), isel, charge(k), kappa(k), eta(k), npart(k), nsh
     allocate(lam(nsh), npr(nsh), orig(3, nsh), ex(0), co(0))
     do s = 1, nsh
        read(u, *) lam(s), npr(s), orig(1, s), orig(2, s), orig(3, s)
        allocate(tmpe(npr(s)), tmpc(npr(s)))
        read(u, *) tmpe
        read(u, *) tmpc
        ex = [ex, tmpe]
        co = [co, tmpc]
        deallocate(tmpe, tmpc)
     end do
     np_total = size(ex)
     call Libint2Interface_constructor(objs(k), ids(k), isel /= 0, lam, npr, orig, ex, co)
     nbas(k) = objs(k)%nbasis
     deallocate(lam, npr, orig, ex, co)
  end do
  close(u)

  ! optional ghost.txt: id of the SCF_GHOST_SPECIES, whose coupling is left
  ! out of the other species' coupling matrices (WaveFunction.f90:1502-1503)
  ghost_id = -1
  inquire(file=trim(dir)//"/ghost.txt", exist=have)
  if (have) then
     open(newunit=u, file=trim(dir)//"/ghost.txt", status="old", action="read")
     read(u, *) ghost_id
     close(u)
  end if

  do k = 1, ns
     allocate(dens(k)%v(nbas(k), nbas(k)))
     call read_matrix(trim(dir)//"/dens_"//itoa(ids(k))//".bin", dens(k)%v)
  end do

  do k = 1, ns
     ! two-particle matrix (direct branch)
     factor = kappa(k) / eta(k)
     call DirectIntegralManager_getDirectIntraRepulsionMatrix(k, "LIBINT", dens(k)%v, G, factor, objs)
     call write_matrix(trim(dir)//"/g_"//itoa(ids(k))//".bin", G)
     if (npart(k) > 1) then
        G = G * charge(k)**2
     else
        G = 0.0_8
     end if
     call write_matrix(trim(dir)//"/two_"//itoa(ids(k))//".bin", G)
     ! coupling matrix (direct branch)
     allocate(coupling(nbas(k), nbas(k)))
     coupling = 0.0_8
     do j = 1, ns
        if (j == k) cycle
        call DirectIntegralManager_getDirectInterRepulsionMatrix(k, j, "LIBINT", dens(j)%v, aux, objs)
        call write_matrix(trim(dir)//"/c_"//itoa(ids(k))//"_"//itoa(ids(j))//".bin", aux)
        aux = aux * charge(k) * charge(j)
        do i = 1, nbas(k)
           do i2 = i, nbas(k)
              aux(i2, i) = aux(i, i2)
           end do
        end do
        if (ids(j) /= ghost_id) coupling = coupling + aux
     end do
     call write_matrix(trim(dir)//"/coupling_"//itoa(ids(k))//".bin", coupling)
     deallocate(coupling)
  end do

  ! optional: Gaussian-geminal builds (g12.txt: term count, exponents,
  ! coefficients) and first-quarter transforms (it.txt: orbital row p, the
  ! density used as the coefficient matrix)
  inquire(file=trim(dir)//"/g12.txt", exist=have)
  if (have) then
     open(newunit=u, file=trim(dir)//"/g12.txt", status="old", action="read")
     read(u, *) ng
     allocate(gexp(ng), gcoef(ng))
     read(u, *) gexp
     read(u, *) gcoef
     close(u)
     do k = 1, ns
        factor = kappa(k) / eta(k)
        call Libint2Interface_computeG12Intraspecies_direct(dens(k)%v, G, factor, gcoef, gexp, objs(k))
        call write_matrix(trim(dir)//"/g12_"//itoa(ids(k))//".bin", G)
        do j = 1, ns
           if (j == k) cycle
           call Libint2Interface_computeG12Interspecies_direct(dens(j)%v, aux, gcoef, gexp, objs(k), objs(j))
           call write_matrix(trim(dir)//"/g12c_"//itoa(ids(k))//"_"//itoa(ids(j))//".bin", aux)
        end do
     end do
  end if
  inquire(file=trim(dir)//"/it.txt", exist=have)
  if (have) then
     open(newunit=u, file=trim(dir)//"/it.txt", status="old", action="read")
     read(u, *) p
     close(u)
     do k = 1, ns
        call Libint2Interface_compute2BodyIntraspecies_direct_IT(dens(k)%v, dens(k)%v, A3, p, objs(k))
        open(newunit=u, file=trim(dir)//"/it_"//itoa(ids(k))//".bin", access="stream", form="unformatted", &
             status="replace", action="write")
        write(u) A3
        close(u)
     end do
  end if

  ! optional all.txt (orbital row p): the integral stores through the
  ! MEMORY-path wrappers -- every species' own, and the inter-species one
  ! from the lower species id -- and the inter-species first-quarter
  ! transform on row p (the density as the coefficient matrix)
  inquire(file=trim(dir)//"/all.txt", exist=have)
  if (have) then
     open(newunit=u, file=trim(dir)//"/all.txt", status="old", action="read")
     read(u, *) p
     close(u)
     do k = 1, ns
        allocate(A4(nbas(k), nbas(k), nbas(k), nbas(k)))
        A4 = 0.0_8
        call Libint2Interface_compute2BodyIntraspecies_direct_all(dens(k)%v, A4, objs(k))
        call write_raw(trim(dir)//"/all_"//itoa(ids(k))//".bin", A4)
        deallocate(A4)
        do j = 1, ns
           if (j == k) cycle
           if (ids(k) < ids(j)) then
              allocate(A4(nbas(j), nbas(j), nbas(k), nbas(k)))
              A4 = 0.0_8
              call Libint2Interface_compute2BodyInterspecies_direct_all(dens(k)%v, A4, objs(k), objs(j))
              call write_raw(trim(dir)//"/allc_"//itoa(ids(k))//"_"//itoa(ids(j))//".bin", A4)
              deallocate(A4)
           end if
           call Libint2Interface_compute2BodyInterspecies_direct_IT(dens(k)%v, dens(k)%v, A3, p, objs(k), objs(j))
           open(newunit=u, file=trim(dir)//"/itc_"//itoa(ids(k))//"_"//itoa(ids(j))//".bin", access="stream", &
                form="unformatted", status="replace", action="write")
           write(u) A3
           close(u)
        end do
     end do
  end if

  do k = 1, ns
     call Libint2Interface_destructor(objs(k))
  end do
  write(*, '(A,I0,A)') "fock_host_driver: ", ns, " species done"

contains

  function itoa(i) result(s)
    integer, intent(in) :: i
    character(len=:), allocatable :: s
    character(len=32) :: buf
    write(buf, '(I0)') i
    s = trim(buf)
  end function itoa

  subroutine read_matrix(path, m)
    character(*), intent(in) :: path
    real(8), intent(out) :: m(:,:)
    integer :: uu
    open(newunit=uu, file=path, access="stream", form="unformatted", status="old", action="read")
    read(uu) m
    close(uu)
  end subroutine read_matrix

  subroutine write_raw(path, a)
    character(*), intent(in) :: path
    real(8), intent(in) :: a(:,:,:,:)
    integer :: uu
    open(newunit=uu, file=path, access="stream", form="unformatted", status="replace", action="write")
    write(uu) a
    close(uu)
  end subroutine write_raw

  subroutine write_matrix(path, m)
    character(*), intent(in) :: path
    real(8), intent(in) :: m(:,:)
    integer :: uu
    open(newunit=uu, file=path, access="stream", form="unformatted", status="replace", action="write")
    write(uu) m
    close(uu)
  end subroutine write_matrix

end program fock_host_driver
