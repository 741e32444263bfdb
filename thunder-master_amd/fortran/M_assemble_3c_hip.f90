! Drop-in replacement of module M_assemble_3c (j.ASSEMBLERS/M_assemble_3c_DOGS.f90)
! whose three-center neutral-atom blocks come from the GPU library.
!
! assemble_vna_3c adds the bcna three-center terms into s%vna(:)%neighbors(:)%block
! as the reference does (M_assemble_3c_DOGS.f90:86-340).  The library keeps the
! running Vna blocks on the GPU (the atom case from assemble_vna_2c plus every
! three-center term, summed per target pair in the reference's triplet order);
! the difference between them and the blocks the session last handed to the
! caller is added into the caller's blocks, so whatever the caller added in
! between is kept.  The 3c terms land only on non-self blocks, which
! assemble_vna_2c leaves at zero, so an untouched caller gets the library's
! values bit for bit; on a block the caller changed, the block's 3c terms are
! added as one sum (the reference adds them one at a time), which can change
! the last bits.  It needs assemble_S, assemble_dipole_z and
! assemble_vna_2c of this session first, as the charged terms use the S and
! dipole_z blocks.
        module M_assemble_3c
        use iso_c_binding
        use fb_hip_c
        use M_fb_hip_session
        use M_assemble_blocks
        use M_species
        use M_configuraciones
        implicit none

        contains

        subroutine assemble_vna_3c (s)
        type(T_structure), target :: s
        call fb_session_sync (s)
        call fb_check (fb_assemble_vna_3c(fbh), fbh, 'assemble_vna_3c')
        call fb_session_fetch (FB_MAT_VNA)
        call fb_session_add_vna (s)
        end subroutine assemble_vna_3c

        end module M_assemble_3c
