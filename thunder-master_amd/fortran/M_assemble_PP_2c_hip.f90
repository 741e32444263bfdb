! Drop-in replacement of module M_assemble_PP_2c (j.ASSEMBLERS/M_assemble_PP_2c.f90)
! whose pseudopotential blocks come from the GPU library.
!
! Same module name, public subroutines and effect on the caller's T_structure:
! assemble_svnl allocates and fills s%svnl(iatom)%neighbors(ineigh)%block
! (norb_mu x norb_PP_nu, one per neighbors_PP entry); assemble_vnl_2c allocates
! s%vnl(iatom)%neighbors(ineigh)%block (norb_mu x norb_nu, one per
! neighbors_PPp entry) and fills the atom and ontop terms;
! destroy_assemble_PP_2c frees both.  The caller's driver_neighbors_PP must
! have run (as for the reference's assemblers): the library builds the same
! lists and stops the run if they differ.
        module M_assemble_PP_2c
        use iso_c_binding
        use fb_hip_c
        use M_fb_hip_session
        use M_assemble_blocks
        use M_species
        use M_configuraciones
        implicit none

        contains

! M_assemble_PP_2c.f90:95-207
        subroutine assemble_svnl (s)
        type(T_structure), target :: s
        call fb_session_sync (s)
        call fb_session_pp (s)
        call fb_check (fb_assemble_svnl(fbh), fbh, 'assemble_svnl')
        allocate (s%svnl(s%natoms))
        call fb_session_ppstore (s, s%svnl, 0, .true.)
        end subroutine assemble_svnl

! M_assemble_PP_2c.f90:230-496
        subroutine assemble_vnl_2c (s)
        type(T_structure), target :: s
        call fb_session_sync (s)
        call fb_session_pp (s)
        call fb_check (fb_assemble_vnl_2c(fbh), fbh, 'assemble_vnl_2c')
        allocate (s%vnl(s%natoms))
        call fb_session_ppstore (s, s%vnl, 1, .true.)
        end subroutine assemble_vnl_2c

! M_assemble_PP_2c.f90:519-560
        subroutine destroy_assemble_PP_2c (s)
        type(T_structure), target :: s
        integer iatom, ineigh
        do iatom = 1, s%natoms
          do ineigh = 1, s%neighbors_PP(iatom)%neighn
            deallocate (s%svnl(iatom)%neighbors(ineigh)%block)
          end do
          deallocate (s%svnl(iatom)%neighbors)
        end do
        deallocate (s%svnl)
        do iatom = 1, s%natoms
          do ineigh = 1, s%neighbors_PPp(iatom)%neighn
            deallocate (s%vnl(iatom)%neighbors(ineigh)%block)
          end do
          deallocate (s%vnl(iatom)%neighbors)
        end do
        deallocate (s%vnl)
        end subroutine destroy_assemble_PP_2c

        end module M_assemble_PP_2c
