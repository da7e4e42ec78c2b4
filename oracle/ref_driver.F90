! Golden-vector driver for the reference CPU path.
!
! This program is test infrastructure.  It is compiled together with the
! reference's own CPU modules (src/half_{kinds,constants,math,types,basis,fft,
! xc,potential,paw,dense_solver,forces,kpoints,energy,chgcar,potcar}.F90, read
! in place from the reference tree by oracle/Makefile) and writes plain
! "key value..." lines that tests/golden/make_golden.py turns into fixtures.
!
! Usage:
!   half_ref_driver solve CHGCAR POTCAR ENCUT lda|pbe NBANDS KX KY KZ VEFF_OUT [uspp]
!   half_ref_driver kat
!   half_ref_driver forces CHGCAR POTCAR ENCUT lda|pbe NBANDS KX KY KZ [uspp]
!   half_ref_driver time CHGCAR POTCAR ENCUT lda|pbe NBANDS
!
! With "uspp" the potential-dependent D correction of src/half_uspp.F90
! (build_uspp_dij_cpu, QDEP) is applied after the projectors are built, as
! src/half_library.F90:277 does for a context created with use_uspp, and the
! per-atom D and dD/dR are printed ("dij", "ddij").
!
! "solve" mirrors the CPU context of src/half_library.F90:51-110 (create) and
! :329-388 (solve_kpoint) plus the Harris-energy assembly of
! app/half_cli.F90:896-920 and the non-local force of src/half_forces.F90:9-73.
! "kat" evaluates the reference routines on the known-answer inputs of
! app/half_energy_check.F90:19-72.
! "forces" mirrors the analytic force assembly of app/half_cli.F90:921-962
! for one k point of weight 1 with zero-temperature occupations: smooth
! output density (src/half_local_forces.F90 accumulate_smooth_density),
! augmentation occupancy and density (src/half_uspp.F90:26-111), Ewald,
! local, non-local (projector + dD/dR parts), NLCC and Harris-correction
! forces.
! "time" runs the reference CPU context's Gamma pipeline once -- Veff
! (build_veff_*), basis, projectors (build_paw_operators) and the dense
! solve (solve_dense_gamma: assembly, overlap spectrum, zhegv) -- and prints
! the wall time of each stage ("time_<stage> seconds", system_clock).
program half_ref_driver
  use half_kinds, only: dp, i32, i64
  use half_types, only: crystal_t, charge_grid_t, plane_wave_basis_t, potcar_t
  use half_chgcar, only: read_chgcar
  use half_potcar, only: read_potcar, validate_potcar_structure
  use half_basis, only: build_plane_wave_basis
  use half_potential, only: build_veff_lda, build_veff_pbe
  use half_paw, only: paw_species_t, build_paw_operators
  use half_dense_solver, only: assemble_dense_gamma, solve_dense_gamma
  use half_forces, only: add_nonlocal_paw_forces
  use half_uspp, only: build_uspp_dij_cpu
  use half_energy, only: compute_occupations, ewald_energy, ewald_forces
  use half_uspp, only: augmentation_occupancy_t, initialize_augmentation_occupancy, &
    accumulate_augmentation_occupancy, add_augmentation_density
  use half_local_forces, only: local_ionic_forces, nlcc_forces, accumulate_smooth_density, &
    harris_correction_forces
  implicit none
  character(len=1024) :: mode
  call get_command_argument(1, mode)
  select case (trim(mode))
  case ('solve')
    call run_solve()
  case ('kat')
    call run_kat()
  case ('forces')
    call run_forces()
  case ('time')
    call run_time()
  case default
    write(0,'(A)') 'usage: half_ref_driver solve|kat ...'
    stop 2
  end select
contains
  subroutine run_solve()
    character(len=1024) :: chg_path, pot_path, arg, veff_path
    type(crystal_t) :: crystal
    type(charge_grid_t) :: charge
    type(potcar_t), allocatable :: potcars(:)
    type(plane_wave_basis_t) :: basis
    type(paw_species_t), allocatable :: paw(:)
    real(dp), allocatable :: veff(:), values(:), occ(:,:), eig2(:,:), charges(:), forces(:,:)
    complex(dp), allocatable :: h(:,:), s(:,:), vectors(:,:), psi(:,:), hpsi(:,:), spsi(:,:)
    real(dp) :: encut, kpt(3), eh, exc, exv, smin, smax, mu, band, entropy, ewald, local_g0, atomic, nelect
    real(dp) :: weights(1)
    integer :: nbands, i, j, it, ion, iat, unit, ib, ig, a, b, al
    integer(i64) :: n
    logical :: pbe, uspp

    call get_command_argument(2, chg_path); call get_command_argument(3, pot_path)
    call get_command_argument(4, arg); read(arg, *) encut
    call get_command_argument(5, arg); pbe = trim(arg) == 'pbe'
    call get_command_argument(6, arg); read(arg, *) nbands
    do i = 1, 3
      call get_command_argument(6 + i, arg); read(arg, *) kpt(i)
    end do
    call get_command_argument(10, veff_path)
    call get_command_argument(11, arg); uspp = trim(arg) == 'uspp'

    call read_chgcar(trim(chg_path), crystal, charge)
    call read_potcar(trim(pot_path), potcars)
    call validate_potcar_structure(potcars, crystal)
    if (pbe) then
      call build_veff_pbe(charge, potcars, crystal, veff, eh, exc, exv)
    else
      call build_veff_lda(charge, potcars, crystal, veff, eh, exc, exv)
    end if
    write(*,'(A,I0)') 'ngrid ', size(veff)
    write(*,'(A,3ES25.17)') 'energy_terms ', eh, exc, exv
    write(*,'(A,ES25.17)') 'volume ', crystal%volume
    write(*,'(A,ES25.17)') 'nelect_grid ', charge%electron_count()
    open(newunit=unit, file=trim(veff_path), access='stream', form='unformatted', status='replace')
    write(unit) veff
    close(unit)

    call build_plane_wave_basis(crystal, charge%shape, encut, kpt, basis)
    n = basis%npw
    write(*,'(A,I0)') 'npw ', n
    call build_paw_operators(potcars, crystal, basis, paw)
    if (uspp) then
      call build_uspp_dij_cpu(veff, charge%shape, potcars, crystal, paw)
      ion = 0
      do it = 1, size(paw)
        do iat = 1, paw(it)%natoms
          ion = ion + 1
          do b = 1, paw(it)%nlm
            do a = 1, paw(it)%nlm
              write(*,'(A,3I5,ES25.17)') 'dij ', ion, a, b, paw(it)%dij_atom(iat, a, b)
              do al = 1, 3
                write(*,'(A,4I5,ES25.17)') 'ddij ', ion, a, b, al, paw(it)%ddij_atom(iat, a, b, al)
              end do
            end do
          end do
        end do
      end do
    end if
    call assemble_dense_gamma(veff, basis, paw, h, s)
    do j = 1, int(n), 41
      do i = 1, int(n), 37
        write(*,'(A,2I7,4ES25.17)') 'hs ', i, j, real(h(i,j)), aimag(h(i,j)), real(s(i,j)), aimag(s(i,j))
      end do
    end do
    do i = 1, int(n)
      write(*,'(A,I7,2ES25.17)') 'hdiag ', i, real(h(i,i)), real(s(i,i))
    end do
    ! deterministic state block and its H/S images (half_apply_hs on the CPU
    ! backend is matmul(h, psi), src/half_library.F90:322-326)
    allocate(psi(n, 3))
    do ib = 1, 3
      do ig = 1, int(n)
        psi(ig, ib) = cmplx(cos(0.37_dp*ig + ib), sin(0.11_dp*ig*ib), dp)
      end do
    end do
    hpsi = matmul(h, psi); spsi = matmul(s, psi)
    do ib = 1, 3
      do ig = 1, int(n)
        write(*,'(A,2I7,4ES25.17)') 'apply ', ig, ib, real(hpsi(ig,ib)), aimag(hpsi(ig,ib)), &
          real(spsi(ig,ib)), aimag(spsi(ig,ib))
      end do
    end do
    deallocate(h, s)

    call solve_dense_gamma(veff, basis, paw, values, smin, smax, vectors)
    write(*,'(A,2ES25.17)') 'overlap ', smin, smax
    do i = 1, nbands
      write(*,'(A,I5,ES25.17)') 'eig ', i, values(i)
    end do

    ! Harris energy with zero-temperature occupations of the lowest NBANDS
    ! (app/half_cli.F90:896-920; use_uspp is off for these inputs)
    nelect = 0
    do it = 1, size(potcars)
      nelect = nelect + potcars(it)%zval*real(crystal%counts(it), dp)
    end do
    weights = 1.0_dp
    allocate(eig2(1, nbands)); eig2(1,:) = values(:nbands)
    call compute_occupations(eig2, weights, nelect, 0.0_dp, occ, mu, band, entropy)
    allocate(charges(crystal%nions)); ion = 0
    do it = 1, size(potcars)
      do iat = 1, crystal%counts(it); ion = ion + 1; charges(ion) = potcars(it)%zval; end do
    end do
    call ewald_energy(crystal, charges, ewald)
    atomic = 0; local_g0 = 0
    do it = 1, size(potcars)
      atomic = atomic + potcars(it)%eatom*real(crystal%counts(it), dp)
      local_g0 = local_g0 + potcars(it)%psp_local(1)*real(crystal%counts(it), dp)
    end do
    local_g0 = local_g0*nelect/crystal%volume
    write(*,'(A,ES25.17)') 'nelect ', nelect
    write(*,'(A,ES25.17)') 'band_energy ', band
    write(*,'(A,ES25.17)') 'fermi ', mu
    write(*,'(A,ES25.17)') 'ewald ', ewald
    write(*,'(A,ES25.17)') 'local_g0 ', local_g0
    write(*,'(A,ES25.17)') 'atomic_reference ', atomic
    write(*,'(A,ES25.17)') 'harris_energy ', band - eh - exv + exc + ewald + local_g0 + atomic
    do i = 1, nbands
      write(*,'(A,I5,ES25.17)') 'occ ', i, occ(1, i)
    end do

    ! non-local projector force for the occupied bands (k weight 1)
    allocate(forces(crystal%nions, 3)); forces = 0
    call add_nonlocal_paw_forces(basis, paw, values(:nbands), occ(1,:), 1.0_dp, vectors(:, :nbands), forces)
    do i = 1, crystal%nions
      write(*,'(A,I5,3ES25.17)') 'force_nl ', i, forces(i, :)
    end do
  end subroutine run_solve

  subroutine run_forces()
    character(len=1024) :: chg_path, pot_path, arg
    type(crystal_t) :: crystal
    type(charge_grid_t) :: charge, rho_out
    type(potcar_t), allocatable :: potcars(:)
    type(plane_wave_basis_t) :: basis
    type(paw_species_t), allocatable :: paw(:)
    type(augmentation_occupancy_t), allocatable :: occupancy(:)
    real(dp), allocatable :: veff(:), vxc(:), values(:), occ(:,:), eig2(:,:), charges(:)
    real(dp), allocatable :: f_ewald(:,:), f_local(:,:), f_nl(:,:), f_proj(:,:), f_aug(:,:), f_nlcc(:,:), f_harris(:,:)
    complex(dp), allocatable :: vectors(:,:)
    real(dp) :: encut, kpt(3), eh, exc, exv, smin, smax, mu, band, entropy, nelect, aug_charge, smooth_sum
    real(dp) :: weights(1)
    integer :: nbands, i, it, ion, iat, a, b
    logical :: pbe, uspp
    call get_command_argument(2, chg_path); call get_command_argument(3, pot_path)
    call get_command_argument(4, arg); read(arg, *) encut
    call get_command_argument(5, arg); pbe = trim(arg) == 'pbe'
    call get_command_argument(6, arg); read(arg, *) nbands
    do i = 1, 3
      call get_command_argument(6 + i, arg); read(arg, *) kpt(i)
    end do
    call get_command_argument(10, arg); uspp = trim(arg) == 'uspp'
    call read_chgcar(trim(chg_path), crystal, charge)
    call read_potcar(trim(pot_path), potcars)
    call validate_potcar_structure(potcars, crystal)
    if (pbe) then
      call build_veff_pbe(charge, potcars, crystal, veff, eh, exc, exv, vxc)
    else
      call build_veff_lda(charge, potcars, crystal, veff, eh, exc, exv, vxc)
    end if
    call build_plane_wave_basis(crystal, charge%shape, encut, kpt, basis)
    call build_paw_operators(potcars, crystal, basis, paw)
    if (uspp) call build_uspp_dij_cpu(veff, charge%shape, potcars, crystal, paw)
    call solve_dense_gamma(veff, basis, paw, values, smin, smax, vectors)
    nelect = 0
    do it = 1, size(potcars)
      nelect = nelect + potcars(it)%zval*real(crystal%counts(it), dp)
    end do
    weights = 1.0_dp
    allocate(eig2(1, nbands)); eig2(1,:) = values(:nbands)
    call compute_occupations(eig2, weights, nelect, 0.0_dp, occ, mu, band, entropy)
    write(*,'(A,I0)') 'npw ', basis%npw
    do i = 1, nbands
      write(*,'(A,I5,2ES25.17)') 'eig_occ ', i, values(i), occ(1, i)
    end do
    ! output density: smooth part from the bands, then the augmentation part
    rho_out%shape = charge%shape
    allocate(rho_out%values(size(charge%values))); rho_out%values = 0
    call accumulate_smooth_density(basis, vectors(:, :nbands), occ(1,:), 1.0_dp, charge%shape, rho_out%values)
    smooth_sum = sum(rho_out%values)
    call initialize_augmentation_occupancy(paw, occupancy)
    call accumulate_augmentation_occupancy(paw, vectors(:, :nbands), occ(1,:), 1.0_dp, occupancy)
    call add_augmentation_density(potcars, crystal, charge%shape, occupancy, rho_out%values, aug_charge)
    write(*,'(A,ES25.17)') 'smooth_sum ', smooth_sum
    write(*,'(A,ES25.17)') 'aug_charge ', aug_charge
    write(*,'(A,ES25.17)') 'rho_out_sum ', sum(rho_out%values)
    do i = 1, size(rho_out%values), 29
      write(*,'(A,I9,ES25.17)') 'rho_out ', i, rho_out%values(i)
    end do
    ion = 0
    do it = 1, size(paw)
      do iat = 1, paw(it)%natoms
        ion = ion + 1
        do b = 1, paw(it)%nlm
          do a = 1, paw(it)%nlm
            write(*,'(A,3I5,ES25.17)') 'aug_occ ', ion, a, b, occupancy(it)%matrix(iat, a, b)
          end do
        end do
      end do
    end do
    allocate(charges(crystal%nions)); ion = 0
    do it = 1, size(potcars)
      do iat = 1, crystal%counts(it); ion = ion + 1; charges(ion) = potcars(it)%zval; end do
    end do
    allocate(f_ewald(crystal%nions,3), f_local(crystal%nions,3), f_nl(crystal%nions,3), f_proj(crystal%nions,3), &
      f_aug(crystal%nions,3), f_nlcc(crystal%nions,3), f_harris(crystal%nions,3))
    f_nl = 0; f_proj = 0; f_aug = 0
    call add_nonlocal_paw_forces(basis, paw, values(:nbands), occ(1,:), 1.0_dp, vectors(:, :nbands), f_nl, &
      f_proj, f_aug)
    call ewald_forces(crystal, charges, f_ewald)
    call local_ionic_forces(rho_out, potcars, crystal, f_local)
    call nlcc_forces(vxc, potcars, crystal, charge%shape, f_nlcc)
    call harris_correction_forces(charge, rho_out, potcars, crystal, pbe, f_harris)
    do i = 1, crystal%nions
      write(*,'(A,I5,3ES25.17)') 'force_ewald ', i, f_ewald(i, :)
      write(*,'(A,I5,3ES25.17)') 'force_local ', i, f_local(i, :)
      write(*,'(A,I5,3ES25.17)') 'force_nonlocal ', i, f_nl(i, :)
      write(*,'(A,I5,3ES25.17)') 'force_projector ', i, f_proj(i, :)
      write(*,'(A,I5,3ES25.17)') 'force_augmentation ', i, f_aug(i, :)
      write(*,'(A,I5,3ES25.17)') 'force_nlcc ', i, f_nlcc(i, :)
      write(*,'(A,I5,3ES25.17)') 'force_harris ', i, f_harris(i, :)
      write(*,'(A,I5,3ES25.17)') 'force_total ', i, f_ewald(i, :) + f_local(i, :) + f_nl(i, :) + f_nlcc(i, :) + &
        f_harris(i, :)
    end do
  end subroutine run_forces

  subroutine run_time()
    character(len=1024) :: chg_path, pot_path, arg
    type(crystal_t) :: crystal
    type(charge_grid_t) :: charge
    type(potcar_t), allocatable :: potcars(:)
    type(plane_wave_basis_t) :: basis
    type(paw_species_t), allocatable :: paw(:)
    real(dp), allocatable :: veff(:), values(:)
    real(dp) :: encut, eh, exc, exv, smin, smax, kpt(3)
    integer :: nbands
    integer(i64) :: c0, c1, c2, c3, c4, rate
    logical :: pbe
    call get_command_argument(2, chg_path); call get_command_argument(3, pot_path)
    call get_command_argument(4, arg); read(arg, *) encut
    call get_command_argument(5, arg); pbe = trim(arg) == 'pbe'
    call get_command_argument(6, arg); read(arg, *) nbands
    kpt = 0
    call read_chgcar(trim(chg_path), crystal, charge)
    call read_potcar(trim(pot_path), potcars)
    call validate_potcar_structure(potcars, crystal)
    call system_clock(c0, rate)
    if (pbe) then
      call build_veff_pbe(charge, potcars, crystal, veff, eh, exc, exv)
    else
      call build_veff_lda(charge, potcars, crystal, veff, eh, exc, exv)
    end if
    call system_clock(c1)
    call build_plane_wave_basis(crystal, charge%shape, encut, kpt, basis)
    call system_clock(c2)
    call build_paw_operators(potcars, crystal, basis, paw)
    call system_clock(c3)
    call solve_dense_gamma(veff, basis, paw, values, smin, smax)
    call system_clock(c4)
    write(*,'(A,I0)') 'npw ', basis%npw
    write(*,'(A,ES25.17)') 'time_veff ', real(c1 - c0, dp)/real(rate, dp)
    write(*,'(A,ES25.17)') 'time_basis ', real(c2 - c1, dp)/real(rate, dp)
    write(*,'(A,ES25.17)') 'time_projectors ', real(c3 - c2, dp)/real(rate, dp)
    write(*,'(A,ES25.17)') 'time_dense_solve ', real(c4 - c3, dp)/real(rate, dp)
    write(*,'(A,ES25.17)') 'eig_first ', values(1)
    write(*,'(A,ES25.17)') 'eig_last ', values(nbands)
  end subroutine run_time

  subroutine run_kat()
    ! Known-answer inputs of app/half_energy_check.F90:19-72.
    type(crystal_t) :: crystal
    type(plane_wave_basis_t) :: basis
    type(paw_species_t), allocatable :: paw(:)
    real(dp) :: eig(2,3), weight(2), mu, band, entropy, e1, e2, analytic(2,3), phase
    real(dp) :: nlforce(1,3), eval(1), focc(1), r0(3)
    complex(dp) :: waves(3,1), projector0(2,3)
    real(dp), allocatable :: occ(:,:)
    integer :: ig, i

    eig = 0.0_dp; eig(:,1:2) = reshape([0.0_dp,1.0_dp,2.0_dp,3.0_dp],[2,2]); weight = [0.5_dp,0.5_dp]
    call compute_occupations(eig(:,1:2), weight, 2.0_dp, 0.0_dp, occ, mu, band, entropy)
    write(*,'(A,4ES25.17,3ES25.17)') 'occ_zero ', occ, mu, band, entropy
    eig = reshape([-1.0_dp,-0.8_dp,0.5_dp,0.7_dp,2.0_dp,2.2_dp],[2,3]); weight = [0.25_dp,0.75_dp]
    call compute_occupations(eig, weight, 2.0_dp, 0.1_dp, occ, mu, band, entropy, -1)
    write(*,'(A,6ES25.17,3ES25.17)') 'occ_fermi ', occ, mu, band, entropy
    call compute_occupations(eig, weight, 2.0_dp, 0.1_dp, occ, mu, band, entropy, 0)
    write(*,'(A,6ES25.17,3ES25.17)') 'occ_gauss ', occ, mu, band, entropy

    crystal%nions = 2; crystal%ntypes = 1; crystal%lattice = 0.0_dp
    crystal%lattice(1,1) = 4.0_dp; crystal%lattice(2,2) = 4.0_dp; crystal%lattice(3,3) = 4.0_dp
    allocate(crystal%counts(1), crystal%species(1), crystal%positions(2,3))
    crystal%counts = 2; crystal%species = 'X'
    crystal%positions(1,:) = [0.07_dp,0.11_dp,0.19_dp]
    crystal%positions(2,:) = [0.31_dp,0.46_dp,0.73_dp]
    call crystal%update_geometry()
    call ewald_energy(crystal, [1.0_dp,1.0_dp], e1, eta=0.35_dp, tolerance=1e-12_dp)
    call ewald_energy(crystal, [1.0_dp,1.0_dp], e2, eta=0.80_dp, tolerance=1e-12_dp)
    write(*,'(A,2ES25.17)') 'ewald_kat ', e1, e2
    call ewald_forces(crystal, [1.0_dp,1.0_dp], analytic, eta=0.35_dp, tolerance=1e-12_dp)
    do i = 1, 2
      write(*,'(A,I3,3ES25.17)') 'ewald_force_kat ', i, analytic(i,:)
    end do

    basis%npw = 3; allocate(basis%qvectors(3,3))
    basis%qvectors = reshape([0.2_dp,0.7_dp,-0.4_dp, -0.3_dp,0.5_dp,0.8_dp, 0.6_dp,-0.2_dp,0.9_dp],[3,3])
    allocate(paw(1)); paw(1)%natoms = 1; paw(1)%nlm = 2
    allocate(paw(1)%projectors(1,2,3), paw(1)%dij(2,2), paw(1)%qij(2,2), paw(1)%dij_atom(1,2,2))
    projector0 = reshape([cmplx(0.4_dp,0.1_dp,dp), cmplx(-0.2_dp,0.3_dp,dp), &
      cmplx(0.1_dp,-0.2_dp,dp), cmplx(0.5_dp,0.4_dp,dp), cmplx(-0.3_dp,0.2_dp,dp), &
      cmplx(0.2_dp,0.6_dp,dp)], [2,3])
    waves(:,1) = [cmplx(0.3_dp,-0.1_dp,dp), cmplx(-0.2_dp,0.4_dp,dp), cmplx(0.5_dp,0.2_dp,dp)]
    paw(1)%dij_atom(1,:,:) = reshape([1.2_dp,0.15_dp,0.15_dp,0.8_dp],[2,2])
    paw(1)%dij = paw(1)%dij_atom(1,:,:); paw(1)%qij = reshape([0.1_dp,0.02_dp,0.02_dp,0.07_dp],[2,2])
    eval = [0.37_dp]; focc = [1.6_dp]; r0 = [0.13_dp,-0.21_dp,0.08_dp]
    do ig = 1, 3
      phase = dot_product(basis%qvectors(ig,:), r0)
      paw(1)%projectors(1,:,ig) = projector0(:,ig)*cmplx(cos(phase), -sin(phase), dp)
    end do
    nlforce = 0.0_dp
    call add_nonlocal_paw_forces(basis, paw, eval, focc, 0.4_dp, waves, nlforce)
    write(*,'(A,3ES25.17)') 'nonlocal_force_kat ', nlforce(1,:)
  end subroutine run_kat
end program half_ref_driver
