! Drop-in replacement of module M_Dassemble_2c (k.DASSEMBLERS/M_Dassemble_2c_DOGS.f90)
! whose two-center force terms come from the GPU library.
!
! Same module name, public subroutines and effect on the caller's T_structure:
! Dassemble_S / _T / _dipole_z allocate and fill
! s%overlap / s%kinetic / s%dipole_z (iatom)%neighbors(ineigh)%Dblock(3, norb_mu,
! norb_nu); Dassemble_vna_2c allocates and fills s%forces(iatom)%vna_atom(3,
! neighn) and %vna_ontop(3, neighn) from the density matrix in s%denmat; the
! destructor frees them.  The matching assemble_* of the drop-in M_assemble_2c
! must have run on the same structure (the library keeps its lists and blocks).
        module M_Dassemble_2c
        use iso_c_binding
        use fb_hip_c
        use M_fb_hip_session
        use M_assemble_blocks
        use M_species
        use M_configuraciones
        implicit none

        contains

! k.DASSEMBLERS/M_Dassemble_2c_DOGS.f90:95-266
        subroutine Dassemble_S (s)
        type(T_structure), target :: s
        call fb_session_sync (s)
        call fb_check (fb_dassemble_S(fbh), fbh, 'Dassemble_S')
        call fb_session_dstore (s, s%overlap, FB_MAT_S)
        end subroutine Dassemble_S

! k.DASSEMBLERS/M_Dassemble_2c_DOGS.f90:292-463
        subroutine Dassemble_T (s)
        type(T_structure), target :: s
        call fb_session_sync (s)
        call fb_check (fb_dassemble_T(fbh), fbh, 'Dassemble_T')
        call fb_session_dstore (s, s%kinetic, FB_MAT_T)
        end subroutine Dassemble_T

! k.DASSEMBLERS/M_Dassemble_2c_DOGS.f90:487-655
        subroutine Dassemble_dipole_z (s)
        type(T_structure), target :: s
        call fb_session_sync (s)
        call fb_check (fb_dassemble_dipole_z(fbh), fbh, 'Dassemble_dipole_z')
        call fb_session_dstore (s, s%dipole_z, FB_MAT_DIP)
        end subroutine Dassemble_dipole_z

! k.DASSEMBLERS/M_Dassemble_2c_DOGS.f90:679-1199: ontop and atom cases of the
! neutral-atom force, contracted with s%denmat on the GPU.  The shell charges
! are read at call time, as the reference reads s%atom(:)%shell(:)%dQ here
! (a charge mixer may have changed them since assemble_vna_2c).
        subroutine Dassemble_vna_2c (s)
        type(T_structure), target :: s
        integer iatom, n
        integer(c_int64_t) :: npairs, ntrip
        integer(c_int) :: mb
        real(c_double), allocatable, target :: fa(:, :), fo(:, :)
        integer p
        call fb_session_sync (s)
        call fb_session_charges (s)
        call fb_session_denmat (s)
        call fb_check (fb_dassemble_vna_2c(fbh), fbh, 'Dassemble_vna_2c')
        call fb_check (fb_get_list_sizes(fbh, 0, npairs, ntrip, mb), fbh, 'fb_get_list_sizes')
        allocate (fa(3, max(npairs, 1_8)), fo(3, max(npairs, 1_8)))
        call fb_check (fb_get_forces(fbh, 0, c_loc(fa), c_loc(fo), c_null_ptr, c_null_ptr,       &
     &    c_null_ptr), fbh, 'fb_get_forces')
        p = 0
        do iatom = 1, s%natoms
          n = s%neighbors(iatom)%neighn
          allocate (s%forces(iatom)%vna_atom(3, n))
          allocate (s%forces(iatom)%vna_ontop(3, n))
          s%forces(iatom)%vna_atom = fa(:, p+1:p+n)
          s%forces(iatom)%vna_ontop = fo(:, p+1:p+n)
          p = p + n
        end do
        deallocate (fa, fo)
        end subroutine Dassemble_vna_2c

! k.DASSEMBLERS/M_Dassemble_2c_DOGS.f90:1223-1262
        subroutine destroy_Dassemble_2c (s)
        type(T_structure), target :: s
        integer iatom, ineigh
        do iatom = 1, s%natoms
          do ineigh = 1, s%neighbors(iatom)%neighn
            deallocate (s%overlap(iatom)%neighbors(ineigh)%Dblock)
            deallocate (s%kinetic(iatom)%neighbors(ineigh)%Dblock)
            deallocate (s%dipole_z(iatom)%neighbors(ineigh)%Dblock)
          end do
          deallocate (s%forces(iatom)%vna_atom)
          deallocate (s%forces(iatom)%vna_ontop)
        end do
        end subroutine destroy_Dassemble_2c

        end module M_Dassemble_2c
