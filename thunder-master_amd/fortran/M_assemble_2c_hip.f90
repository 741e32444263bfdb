! Drop-in replacement of module M_assemble_2c (j.ASSEMBLERS/M_assemble_2c_DOGS.f90)
! whose two-center blocks come from the GPU library.
!
! Same module name, same public subroutines and the same effect on the
! caller's T_structure: s%overlap, s%kinetic, s%dipole_z and s%vna are
! allocated per atom and per neighbor (norb_mu x norb_nu blocks in
! s%neighbors order) exactly where the reference allocates them, and the
! destroy_* routines free them the same way.
        module M_assemble_2c
        use iso_c_binding
        use fb_hip_c
        use M_fb_hip_session
        use M_assemble_blocks
        use M_species
        use M_configuraciones
        implicit none

        contains

! j.ASSEMBLERS/M_assemble_2c_DOGS.f90:87-209
        subroutine assemble_S (s)
        type(T_structure), target :: s
        call fb_session_sync (s)
        call fb_check (fb_assemble_S(fbh), fbh, 'assemble_S')
        call fb_session_fetch (FB_MAT_S)
        allocate (s%overlap(s%natoms))
        call fb_session_store (s, s%overlap, 0, .true.)
        end subroutine assemble_S

! j.ASSEMBLERS/M_assemble_2c_DOGS.f90:232-353
        subroutine assemble_T (s)
        type(T_structure), target :: s
        call fb_session_sync (s)
        call fb_check (fb_assemble_T(fbh), fbh, 'assemble_T')
        call fb_session_fetch (FB_MAT_T)
        allocate (s%kinetic(s%natoms))
        call fb_session_store (s, s%kinetic, 0, .true.)
        end subroutine assemble_T

! j.ASSEMBLERS/M_assemble_2c_DOGS.f90:376-499
        subroutine assemble_dipole_z (s)
        type(T_structure), target :: s
        call fb_session_sync (s)
        call fb_check (fb_assemble_dipole_z(fbh), fbh, 'assemble_dipole_z')
        call fb_session_fetch (FB_MAT_DIP)
        allocate (s%dipole_z(s%natoms))
        call fb_session_store (s, s%dipole_z, 0, .true.)
        end subroutine assemble_dipole_z

! j.ASSEMBLERS/M_assemble_2c_DOGS.f90:523-829: ontop blocks (blocko) and the
! atom case (block of the self pair, zero elsewhere until assemble_vna_3c).
! The shell charges are read from s%atom(:)%shell(:)%dQ at this call.
        subroutine assemble_vna_2c (s)
        type(T_structure), target :: s
        call fb_session_sync (s)
        call fb_session_charges (s)
        call fb_check (fb_assemble_vna_2c(fbh), fbh, 'assemble_vna_2c')
        allocate (s%vna(s%natoms))
        call fb_session_fetch (FB_MAT_VNA)
        call fb_session_store (s, s%vna, 0, .true.)
        call fb_session_keep_vna ()
        call fb_session_fetch (FB_MAT_VNAO)
        call fb_session_store (s, s%vna, 1, .true.)
        end subroutine assemble_vna_2c

! j.ASSEMBLERS/M_assemble_2c_DOGS.f90:853-890
        subroutine destroy_assemble_vna (s)
        type(T_structure), target :: s
        integer iatom, ineigh
        do iatom = 1, s%natoms
          do ineigh = 1, s%neighbors(iatom)%neighn
            deallocate (s%vna(iatom)%neighbors(ineigh)%block)
            deallocate (s%vna(iatom)%neighbors(ineigh)%blocko)
          end do
          deallocate (s%vna(iatom)%neighbors)
        end do
        deallocate (s%vna)
        end subroutine destroy_assemble_vna

! j.ASSEMBLERS/M_assemble_2c_DOGS.f90:914-956
        subroutine destroy_assemble_2c (s)
        type(T_structure), target :: s
        integer iatom, ineigh
        do iatom = 1, s%natoms
          do ineigh = 1, s%neighbors(iatom)%neighn
            deallocate (s%overlap(iatom)%neighbors(ineigh)%block)
            deallocate (s%kinetic(iatom)%neighbors(ineigh)%block)
            deallocate (s%dipole_z(iatom)%neighbors(ineigh)%block)
          end do
          deallocate (s%overlap(iatom)%neighbors)
          deallocate (s%kinetic(iatom)%neighbors)
          deallocate (s%dipole_z(iatom)%neighbors)
        end do
        deallocate (s%overlap)
        deallocate (s%kinetic)
        deallocate (s%dipole_z)
        end subroutine destroy_assemble_2c

        end module M_assemble_2c
