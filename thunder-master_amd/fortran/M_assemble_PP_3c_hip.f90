! Drop-in replacement of module M_assemble_PP_3c (j.ASSEMBLERS/M_assemble_PP_3c.f90).
!
! assemble_vnl_3c adds the three-center pseudopotential terms into
! s%vnl(:)%neighbors(:)%block as the reference does (M_assemble_PP_3c.f90:94-237);
! the library keeps the running vnl blocks on the GPU (the two-center terms of
! assemble_vnl_2c plus every triplet, in the reference's order), so the result
! is copied over the caller's blocks.  mpairnay, the module's public neighbour
! search, is kept with the reference's interface and meaning.
        module M_assemble_PP_3c
        use iso_c_binding
        use fb_hip_c
        use M_fb_hip_session
        use M_assemble_blocks
        use M_species
        use M_configuraciones
        implicit none

        contains

        subroutine assemble_vnl_3c (s)
        type(T_structure), target :: s
        call fb_session_sync (s)
        call fb_session_pp (s)
        call fb_check (fb_assemble_vnl_3c(fbh), fbh, 'assemble_vnl_3c')
        call fb_session_ppstore (s, s%vnl, 1, .false.)
        end subroutine assemble_vnl_3c

! The neighbors_PP entry of iatom that is jatom at the offset r12 from iatom's
! own position (M_assemble_PP_3c.f90:266-364); exactly one entry must match
! within 1e-4, otherwise the run stops, as in the reference.
        subroutine mpairnay (t, iatom, jatom, r12, mpair)
        type(T_structure), target :: t
        integer, intent(in) :: iatom, jatom
        real, intent(in), dimension (3) :: r12
        integer, intent(out) :: mpair
        integer ineigh, imatch, mbeta
        real, dimension (3) :: d
        mpair = 0
        imatch = 0
        do ineigh = 1, t%neighbors_PP(iatom)%neighn
          if (t%neighbors_PP(iatom)%neigh_j(ineigh) .ne. jatom) cycle
          mbeta = t%neighbors_PP(iatom)%neigh_b(ineigh)
          d = r12 - ((t%atom(jatom)%ratom + t%xl(mbeta)%a) - t%atom(iatom)%ratom)
          if (sqrt(d(1)**2 + d(2)**2 + d(3)**2) .lt. 0.0001d0) then
            imatch = imatch + 1
            mpair = ineigh
          end if
        end do
        if (imatch .ne. 1) then
          write (*,*) ' mpairnay: ', imatch, ' matches of atom ', jatom, ' among the PP neighbors of ', iatom
          error stop ' fireball_hip failure '
        end if
        end subroutine mpairnay

        end module M_assemble_PP_3c
