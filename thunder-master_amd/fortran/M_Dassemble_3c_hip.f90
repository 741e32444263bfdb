! Drop-in replacement of module M_Dassemble_3c (k.DASSEMBLERS/M_Dassemble_3c_DOGS.f90)
! whose three-center neutral-atom forces come from the GPU library.
!
! Dassemble_vna_3c adds the library's per-atom f3naa (on ialpha), f3nab (on
! iatom) and f3nac (on jatom) to s%forces(:), as the reference accumulates them
! (M_Dassemble_3c_DOGS.f90:136-793), from the density matrix in s%denmat and
! the shell charges s%atom(:)%shell(:)%dQ as they are at call time.  It needs
! assemble_S, assemble_dipole_z, Dassemble_S and Dassemble_dipole_z of this
! session first (the charged terms use their blocks and Dblocks).
        module M_Dassemble_3c
        use iso_c_binding
        use fb_hip_c
        use M_fb_hip_session
        use M_assemble_blocks
        use M_species
        use M_configuraciones
        implicit none

        contains

        subroutine Dassemble_vna_3c (s)
        type(T_structure), target :: s
        real(c_double), allocatable, target :: fa(:, :), fb(:, :), fc(:, :)
        integer iatom
        call fb_session_sync (s)
        call fb_session_charges (s)
        call fb_session_denmat (s)
        call fb_check (fb_dassemble_vna_3c(fbh), fbh, 'Dassemble_vna_3c')
        allocate (fa(3, s%natoms), fb(3, s%natoms), fc(3, s%natoms))
        call fb_check (fb_get_forces(fbh, 0, c_null_ptr, c_null_ptr, c_loc(fa), c_loc(fb),       &
     &    c_loc(fc)), fbh, 'fb_get_forces')
        do iatom = 1, s%natoms
          s%forces(iatom)%f3naa = s%forces(iatom)%f3naa + fa(:, iatom)
          s%forces(iatom)%f3nab = s%forces(iatom)%f3nab + fb(:, iatom)
          s%forces(iatom)%f3nac = s%forces(iatom)%f3nac + fc(:, iatom)
        end do
        deallocate (fa, fb, fc)
        end subroutine Dassemble_vna_3c

        end module M_Dassemble_3c
