! ISO_C_BINDING interfaces of the C ABI in include/fireball_hip.h.
!
! Only the entry points the drop-in assembler modules use are bound here; the
! argument order and meaning are those of the header.
        module fb_hip_c
        use iso_c_binding
        implicit none

        integer(c_int), parameter :: FB_OK = 0
        integer(c_int), parameter :: FB_MAT_S = 0, FB_MAT_T = 1, FB_MAT_DIP = 2
        integer(c_int), parameter :: FB_MAT_VNA = 3, FB_MAT_VNAO = 4
        integer(c_int), parameter :: FB_MAT_RHO_IN = 5, FB_MAT_RHO_BOND = 6
        integer(c_int), parameter :: FB_MAT_RHOW_IN = 7, FB_MAT_RHOW_BOND = 8
        integer(c_int), parameter :: FB_PREC_FP64 = 0, FB_PREC_MIXED = 1
        integer, parameter :: FB_MAXSH = 4

        interface
          integer(c_int) function fb_create (ctx, device) bind(c, name='fb_create')
            import :: c_ptr, c_int
            type(c_ptr) :: ctx
            integer(c_int), value :: device
          end function

          type(c_ptr) function fb_last_error (ctx) bind(c, name='fb_last_error')
            import :: c_ptr
            type(c_ptr), value :: ctx
          end function

          integer(c_int) function fb_set_precision (ctx, mode) bind(c, name='fb_set_precision')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: mode
          end function

          integer(c_int) function fb_set_species (ctx, ns, nZ, nssh, lssh, rcA)   &
     &      bind(c, name='fb_set_species')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: ns
            integer(c_int) :: nZ(*), nssh(*), lssh(*)
            real(c_double) :: rcA(*)
          end function

          integer(c_int) function fb_set_table_2c (ctx, is, js, itype, isub, nME, nx,  &
     &      xmax, mu, nu, mvalue, fdata) bind(c, name='fb_set_table_2c')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: is, js, itype, isub, nME, nx
            real(c_double), value :: xmax
            integer(c_int) :: mu(*), nu(*), mvalue(*)
            real(c_double) :: fdata(*)
          end function

          integer(c_int) function fb_set_table_3c (ctx, is, js, ks, itype, isub, itheta,&
     &      nME, nx, xmax, ny, ymax, mu, nu, mvalue, fdata) bind(c, name='fb_set_table_3c')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: is, js, ks, itype, isub, itheta, nME, nx, ny
            real(c_double), value :: xmax, ymax
            integer(c_int) :: mu(*), nu(*), mvalue(*)
            real(c_double) :: fdata(*)
          end function

          integer(c_int) function fb_commit_tables (ctx) bind(c, name='fb_commit_tables')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function

          integer(c_int) function fb_clear_structures (ctx) bind(c, name='fb_clear_structures')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function

          integer(c_int) function fb_add_structure (ctx, natoms, ratom, imass, lattice,   &
     &      icluster, dQ, istruct) bind(c, name='fb_add_structure')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: natoms, icluster
            real(c_double) :: ratom(*), lattice(*), dQ(*)
            integer(c_int) :: imass(*), istruct
          end function

          integer(c_int) function fb_set_positions (ctx, is, ratom) bind(c, name='fb_set_positions')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: is
            real(c_double) :: ratom(*)
          end function

          integer(c_int) function fb_set_charges (ctx, is, dQ) bind(c, name='fb_set_charges')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: is
            real(c_double) :: dQ(*)
          end function

          integer(c_int) function fb_get_neighbors (ctx, is, neighn, neigh_j, neigh_b,    &
     &      neigh_back, neigh_self) bind(c, name='fb_get_neighbors')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: is
            integer(c_int) :: neighn(*), neigh_j(*), neigh_b(*), neigh_back(*), neigh_self(*)
          end function

          integer(c_int) function fb_get_common (ctx, is, ncommon, ic_j, ic_b, jc_j, jc_b,  &
     &      neigh_common) bind(c, name='fb_get_common')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: is
            integer(c_int) :: ncommon(*), ic_j(*), ic_b(*), jc_j(*), jc_b(*), neigh_common(*)
          end function

          integer(c_int) function fb_build_lists (ctx, npairs, ntrip) bind(c, name='fb_build_lists')
            import :: c_ptr, c_int, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t) :: npairs, ntrip
          end function

          integer(c_int) function fb_get_block_len (ctx, is, blen) bind(c, name='fb_get_block_len')
            import :: c_ptr, c_int, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: is
            integer(c_int64_t) :: blen
          end function

          integer(c_int) function fb_get_blocks (ctx, is, matrix, out) bind(c, name='fb_get_blocks')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: is, matrix
            real(c_double) :: out(*)
          end function

          integer(c_int) function fb_assemble_S (ctx) bind(c, name='fb_assemble_S')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function
          integer(c_int) function fb_assemble_T (ctx) bind(c, name='fb_assemble_T')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function
          integer(c_int) function fb_assemble_dipole_z (ctx) bind(c, name='fb_assemble_dipole_z')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function
          integer(c_int) function fb_assemble_vna_2c (ctx) bind(c, name='fb_assemble_vna_2c')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function
          integer(c_int) function fb_assemble_vna_3c (ctx) bind(c, name='fb_assemble_vna_3c')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function

          integer(c_int) function fb_get_list_sizes (ctx, is, npairs, ntrip, mbeta_max)           &
     &        bind(c, name='fb_get_list_sizes')
            import :: c_ptr, c_int, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: is
            integer(c_int64_t) :: npairs, ntrip
            integer(c_int) :: mbeta_max
          end function

          integer(c_int) function fb_set_denmat (ctx, is, rho) bind(c, name='fb_set_denmat')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: is
            real(c_double) :: rho(*)
          end function
          integer(c_int) function fb_dassemble_S (ctx) bind(c, name='fb_dassemble_S')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function
          integer(c_int) function fb_dassemble_T (ctx) bind(c, name='fb_dassemble_T')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function
          integer(c_int) function fb_dassemble_dipole_z (ctx) bind(c, name='fb_dassemble_dipole_z')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function
          integer(c_int) function fb_dassemble_vna_2c (ctx) bind(c, name='fb_dassemble_vna_2c')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function
          integer(c_int) function fb_dassemble_vna_3c (ctx) bind(c, name='fb_dassemble_vna_3c')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function
          integer(c_int) function fb_get_dblocks (ctx, is, matrix, out) bind(c, name='fb_get_dblocks')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: is, matrix
            real(c_double) :: out(*)
          end function
          integer(c_int) function fb_get_forces (ctx, is, vna_atom, vna_ontop, f3naa, f3nab, f3nac) &
     &        bind(c, name='fb_get_forces')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: is
            type(c_ptr), value :: vna_atom, vna_ontop, f3naa, f3nab, f3nac
          end function
          integer(c_int) function fb_set_qin (ctx, is, qin) bind(c, name='fb_set_qin')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: is
            real(c_double) :: qin(*)
          end function
          integer(c_int) function fb_assemble_rho_2c (ctx) bind(c, name='fb_assemble_rho_2c')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function
          integer(c_int) function fb_assemble_rho_3c (ctx) bind(c, name='fb_assemble_rho_3c')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function
          integer(c_int) function fb_assemble_rho_weighted_2c (ctx) bind(c, name='fb_assemble_rho_weighted_2c')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function
          integer(c_int) function fb_assemble_rho_weighted_3c (ctx) bind(c, name='fb_assemble_rho_weighted_3c')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function
          integer(c_int) function fb_get_shell_block_len (ctx, is, wlen) bind(c, name='fb_get_shell_block_len')
            import :: c_ptr, c_int, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: is
            integer(c_int64_t) :: wlen
          end function
          integer(c_int) function fb_set_species_pp (ctx, nssh_pp, lssh_pp, rcutoff_pp, cl)       &
     &        bind(c, name='fb_set_species_pp')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int) :: nssh_pp(*), lssh_pp(*)
            real(c_double) :: rcutoff_pp(*), cl(*)
          end function
          integer(c_int) function fb_build_lists_pp (ctx, npp, nppp, ntrip) bind(c, name='fb_build_lists_pp')
            import :: c_ptr, c_int, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t) :: npp, nppp, ntrip
          end function
          integer(c_int) function fb_get_pp_list_sizes (ctx, is, sizes) bind(c, name='fb_get_pp_list_sizes')
            import :: c_ptr, c_int, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: is
            integer(c_int64_t) :: sizes(6)
          end function
          integer(c_int) function fb_get_pp_lists (ctx, is, which, n, a0, a1, a2, a3, a4, self)    &
     &        bind(c, name='fb_get_pp_lists')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: is, which
            type(c_ptr), value :: n, a0, a1, a2, a3, a4, self
          end function
          integer(c_int) function fb_assemble_svnl (ctx) bind(c, name='fb_assemble_svnl')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function
          integer(c_int) function fb_assemble_vnl_2c (ctx) bind(c, name='fb_assemble_vnl_2c')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function
          integer(c_int) function fb_assemble_vnl_3c (ctx) bind(c, name='fb_assemble_vnl_3c')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
          end function
          integer(c_int) function fb_get_pp_blocks (ctx, is, which, out) bind(c, name='fb_get_pp_blocks')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: is, which
            real(c_double) :: out(*)
          end function
        end interface

        contains

! stop the run with the library's message and a nonzero exit status
        subroutine fb_check (rc, ctx, where)
        integer(c_int), intent(in) :: rc
        type(c_ptr), intent(in) :: ctx
        character(len=*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        type(c_ptr) :: p
        integer i
        if (rc .eq. FB_OK) return
        p = fb_last_error(ctx)
        write (*,'(A,A,A,I4)') ' fireball_hip error in ', where, ': status ', rc
        if (c_associated(p)) then
          call c_f_pointer(p, msg, [512])
          do i = 1, 512
            if (msg(i) .eq. c_null_char) exit
            write (*,'(A)', advance='no') msg(i)
          end do
          write (*,*)
        end if
        error stop ' fireball_hip failure '
        end subroutine fb_check

        end module fb_hip_c
