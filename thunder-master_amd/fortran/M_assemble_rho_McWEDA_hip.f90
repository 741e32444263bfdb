! Drop-in replacement of module M_assemble_rho_McWEDA
! (j.ASSEMBLERS/M_assemble_rho_McWEDA_DOGS.f90) whose density blocks come from
! the GPU library.
!
! Same module name, public subroutines and effect on the caller's T_structure:
! assemble_rho_2c allocates and fills s%rho_in and s%rho_bond (orbital blocks),
! assemble_rho_weighted_2c s%rho_in_weighted and s%rho_bond_weighted (shell
! blocks nssh_mu x nssh_nu); assemble_rho_3c and assemble_rho_weighted_3c add
! the three-center terms into s%rho_in and s%rho_in_weighted;
! destroy_assemble_rho frees all four.  The input shell charges are read from
! s%atom(:)%shell(:)%Qin at every call, as the reference reads them.  The
! library keeps the running density blocks on the GPU, so the three-center
! calls copy the GPU's sums over the caller's blocks (the caller must not
! change them between the two-center and three-center calls).
        module M_assemble_rho_McWEDA
        use iso_c_binding
        use fb_hip_c
        use M_fb_hip_session
        use M_assemble_blocks
        use M_species
        use M_configuraciones
        implicit none

        contains

! M_assemble_rho_McWEDA_DOGS.f90:90-341
        subroutine assemble_rho_2c (s)
        type(T_structure), target :: s
        call fb_session_sync (s)
        call fb_session_qin (s)
        call fb_check (fb_assemble_rho_2c(fbh), fbh, 'assemble_rho_2c')
        allocate (s%rho_in(s%natoms))
        allocate (s%rho_bond(s%natoms))
        call fb_session_fetch (FB_MAT_RHO_IN)
        call fb_session_store (s, s%rho_in, 0, .true.)
        call fb_session_fetch (FB_MAT_RHO_BOND)
        call fb_session_store (s, s%rho_bond, 0, .true.)
        end subroutine assemble_rho_2c

! M_assemble_rho_McWEDA_DOGS.f90:362-511
        subroutine assemble_rho_3c (s)
        type(T_structure), target :: s
        call fb_session_sync (s)
        call fb_session_qin (s)
        call fb_check (fb_assemble_rho_3c(fbh), fbh, 'assemble_rho_3c')
        call fb_session_fetch (FB_MAT_RHO_IN)
        call fb_session_store (s, s%rho_in, 0, .false.)
        end subroutine assemble_rho_3c

! M_assemble_rho_McWEDA_DOGS.f90:534-762
        subroutine assemble_rho_weighted_2c (s)
        type(T_structure), target :: s
        call fb_session_sync (s)
        call fb_session_qin (s)
        call fb_check (fb_assemble_rho_weighted_2c(fbh), fbh, 'assemble_rho_weighted_2c')
        allocate (s%rho_in_weighted(s%natoms))
        allocate (s%rho_bond_weighted(s%natoms))
        call fb_session_wstore (s, s%rho_in_weighted, FB_MAT_RHOW_IN, .true.)
        call fb_session_wstore (s, s%rho_bond_weighted, FB_MAT_RHOW_BOND, .true.)
        end subroutine assemble_rho_weighted_2c

! M_assemble_rho_McWEDA_DOGS.f90:783-924
        subroutine assemble_rho_weighted_3c (s)
        type(T_structure), target :: s
        call fb_session_sync (s)
        call fb_session_qin (s)
        call fb_check (fb_assemble_rho_weighted_3c(fbh), fbh, 'assemble_rho_weighted_3c')
        call fb_session_wstore (s, s%rho_in_weighted, FB_MAT_RHOW_IN, .false.)
        end subroutine assemble_rho_weighted_3c

! M_assemble_rho_McWEDA_DOGS.f90:949-1000
        subroutine destroy_assemble_rho (s)
        type(T_structure), target :: s
        integer iatom, ineigh
        do iatom = 1, s%natoms
          do ineigh = 1, s%neighbors(iatom)%neighn
            deallocate (s%rho_in_weighted(iatom)%neighbors(ineigh)%block)
            deallocate (s%rho_bond_weighted(iatom)%neighbors(ineigh)%block)
          end do
          deallocate (s%rho_in_weighted(iatom)%neighbors)
          deallocate (s%rho_bond_weighted(iatom)%neighbors)
        end do
        deallocate (s%rho_in_weighted)
        deallocate (s%rho_bond_weighted)
        do iatom = 1, s%natoms
          do ineigh = 1, s%neighbors(iatom)%neighn
            deallocate (s%rho_in(iatom)%neighbors(ineigh)%block)
            deallocate (s%rho_bond(iatom)%neighbors(ineigh)%block)
          end do
          deallocate (s%rho_in(iatom)%neighbors)
          deallocate (s%rho_bond(iatom)%neighbors)
        end do
        deallocate (s%rho_in)
        deallocate (s%rho_bond)
        end subroutine destroy_assemble_rho

        end module M_assemble_rho_McWEDA
