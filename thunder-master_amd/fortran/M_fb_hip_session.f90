! GPU session shared by the drop-in M_assemble_2c / M_assemble_3c modules.
!
! One library context per process.  On first use it takes the species from
! M_species (species(:)%shell(:)%lssh, %rcutoffA) and every Fdata cell of the
! assembled interactions from Fdata_bundle_2c / Fdata_bundle_3c
! (e.FDATA/M_Fdata_2c.f90:95-110, e.FDATA/M_Fdata_3c.f90:88-105), so the
! caller's read_Fdata_2c / read_Fdata_3c remain the only table reader.
!
! Per call it mirrors the caller's T_structure into the library: a change of
! natoms, species, lattice or icluster re-adds the structure, a change of
! positions only replaces them, and either rebuilds the lists on the GPU.  The
! rebuilt lists must equal the caller's s%neighbors (driver_neighbors,
! c.SYSTEM/M_neighbors.f90:93-141), pair for pair and triplet for triplet:
! blocks are returned in that pair order, so a mismatch stops the run.
!
! Environment: FIREBALL_HIP_DEVICE (default 0), FIREBALL_HIP_PRECISION
! ("fp64", default, or "mixed").
        module M_fb_hip_session
        use iso_c_binding
        use fb_hip_c
        use M_assemble_blocks
        use M_species
        use M_configuraciones
        use M_Fdata_2c
        use M_Fdata_3c
        implicit none

        type(c_ptr), save :: fbh = c_null_ptr
        integer, save :: c_natoms = -1, c_icluster = -1
        integer(c_int), allocatable, save :: c_imass(:)
        real(c_double), allocatable, save :: c_ratom(:, :)
        real(c_double), save :: c_lattice(3, 3)
        logical, save :: have_lists = .false.
        logical, save :: have_pp_lists = .false.
        integer(c_int64_t), save :: npairs_all = 0, ntrip_all = 0
        real(c_double), allocatable, save :: fb_buf(:)
        ! the Vna blocks (%block) as last handed to the caller
        real(c_double), allocatable, save :: vna_last(:)

        ! 2c interactions of the assemblers (include/interactions_2c.h:5-25):
        ! H and S, then the McWEDA densities rho_* and rhoS_*
        integer, parameter :: n_itype2c = 13
        integer, parameter :: itype2c(n_itype2c) = [1, 2, 3, 4, 5, 6, 10, 15, 16, 17, 19, 20, 21]
        ! 3c: rho_3c, rhoS_3c and bcna (include/interactions_3c.h:6-8) over five
        ! Legendre angles
        integer, parameter :: n_itype3c = 3
        integer, parameter :: itype3c(n_itype3c) = [1, 2, 3]
        integer, parameter :: ntheta3c = 5

        contains

        subroutine fb_session_open ()
        integer(c_int) :: dev, prec
        integer ispecies, issh, nsp, is, js, ks, k, itype, isub, itheta, idx, st
        integer(c_int), allocatable :: nZ(:), nssh(:), lssh(:, :)
        real(c_double), allocatable :: rcA(:, :)
        character(len=32) :: val
        type(T_Fdata_cell_2c), pointer :: c2
        type(T_Fdata_cell_3c), pointer :: c3

        if (c_associated(fbh)) return
        dev = 0
        call get_environment_variable ('FIREBALL_HIP_DEVICE', val, status = st)
        if (st .eq. 0) read (val, *) dev
        call fb_check (fb_create(fbh, dev), fbh, 'fb_create')
        prec = FB_PREC_FP64
        call get_environment_variable ('FIREBALL_HIP_PRECISION', val, status = st)
        if (st .eq. 0 .and. trim(val) .eq. 'mixed') prec = FB_PREC_MIXED
        call fb_check (fb_set_precision(fbh, prec), fbh, 'fb_set_precision')

        nsp = size(species)
        allocate (nZ(nsp), nssh(nsp), lssh(FB_MAXSH, nsp), rcA(FB_MAXSH, nsp))
        lssh = 0
        rcA = 0.0d0
        do ispecies = 1, nsp
          nZ(ispecies) = species(ispecies)%nZ
          nssh(ispecies) = species(ispecies)%nssh
          if (nssh(ispecies) .gt. FB_MAXSH) error stop ' fireball_hip: more than 4 shells '
          do issh = 1, species(ispecies)%nssh
            lssh(issh, ispecies) = species(ispecies)%shell(issh)%lssh
            rcA(issh, ispecies) = species(ispecies)%shell(issh)%rcutoffA
          end do
        end do
        call fb_check (fb_set_species(fbh, nsp, nZ, nssh, lssh, rcA), fbh, 'fb_set_species')
        call set_species_pp (nsp)

        do is = 1, nsp
          do js = 1, nsp
            do k = 1, n_itype2c
              itype = itype2c(k)
              do isub = 0, ubound(Fdata_bundle_2c(is, js)%index_2c, 2)
                idx = Fdata_bundle_2c(is, js)%index_2c(itype, isub)
                if (idx .le. 0) cycle
                c2 => Fdata_bundle_2c(is, js)%Fdata_cell_2c(idx)
                call fb_check (fb_set_table_2c(fbh, is, js, itype, isub, c2%nME, c2%nx,    &
     &            c2%xmax, c2%mu_2c, c2%nu_2c, c2%mvalue_2c, c2%Fdata_2c), fbh,            &
     &            'fb_set_table_2c')
              end do
            end do
            do ks = 1, nsp
              do k = 1, n_itype3c
                itype = itype3c(k)
                do isub = 0, species(ks)%nssh
                  do itheta = 1, ntheta3c
                    idx = Fdata_bundle_3c(is, js, ks)%index_3c(itype, isub, itheta)
                    if (idx .le. 0) cycle
                    c3 => Fdata_bundle_3c(is, js, ks)%Fdata_cell_3c(idx)
                    call fb_check (fb_set_table_3c(fbh, is, js, ks, itype, isub,           &
     &                itheta, c3%nME, c3%nx, c3%xmax, c3%ny, c3%ymax, c3%mu_3c, c3%nu_3c,  &
     &                c3%mvalue_3c, c3%Fdata_3c), fbh, 'fb_set_table_3c')
                  end do
                end do
              end do
            end do
          end do
        end do
        call fb_check (fb_commit_tables(fbh), fbh, 'fb_commit_tables')
        deallocate (nZ, nssh, lssh, rcA)
        end subroutine fb_session_open

! Mirror the caller's structure; rebuild and check the lists when it changed.
        subroutine fb_session_sync (s)
        type(T_structure), target :: s
        if (fb_session_build(s)) call check_lists (s)
        end subroutine fb_session_sync

! Mirror the caller's structure and rebuild the lists on the GPU when it
! changed; .true. when they were rebuilt.
        logical function fb_session_build (s) result (rebuilt)
        type(T_structure), target :: s
        integer iatom, i
        logical same_atoms, same_pos
        integer(c_int) :: idx
        real(c_double), allocatable :: r(:, :), dq(:, :)
        integer(c_int), allocatable :: im(:)

        call fb_session_open ()
        allocate (r(3, s%natoms), im(s%natoms))
        do iatom = 1, s%natoms
          r(:, iatom) = s%atom(iatom)%ratom
          im(iatom) = s%atom(iatom)%imass
        end do
        same_atoms = s%natoms .eq. c_natoms .and. s%icluster .eq. c_icluster
        if (same_atoms) then
          same_atoms = all(im .eq. c_imass)
          do i = 1, 3
            same_atoms = same_atoms .and. all(s%lattice(i)%a .eq. c_lattice(:, i))
          end do
        end if
        same_pos = .false.
        if (same_atoms) same_pos = all(r .eq. c_ratom)
        rebuilt = .false.
        if (same_atoms .and. same_pos .and. have_lists) then
          deallocate (r, im)
          return
        end if

        if (.not. same_atoms) then
          if (allocated(c_imass)) deallocate (c_imass, c_ratom)
          allocate (c_imass(s%natoms), c_ratom(3, s%natoms))
          c_natoms = s%natoms
          c_icluster = s%icluster
          c_imass = im
          do i = 1, 3
            c_lattice(:, i) = s%lattice(i)%a
          end do
          allocate (dq(FB_MAXSH, s%natoms))
          call shell_charges (s, dq)
          call fb_check (fb_clear_structures(fbh), fbh, 'fb_clear_structures')
          call fb_check (fb_add_structure(fbh, c_natoms, r, c_imass, c_lattice,          &
     &      c_icluster, dq, idx), fbh, 'fb_add_structure')
          deallocate (dq)
        else
          call fb_check (fb_set_positions(fbh, 0, r), fbh, 'fb_set_positions')
        end if
        c_ratom = r
        call fb_check (fb_build_lists(fbh, npairs_all, ntrip_all), fbh, 'fb_build_lists')
        have_lists = .true.
        have_pp_lists = .false.
        rebuilt = .true.
        deallocate (r, im)
        end function fb_session_build

! Pseudopotential shells of the species (species(:)%nssh_PP, shell_PP(:)%lssh,
! %rcutoff_PP and the Kleinman-Bylander shell_PP(:)%cl of clPP.ZZ.dat), when
! any species has them.
        subroutine set_species_pp (nsp)
        integer, intent(in) :: nsp
        integer ispecies, issh
        integer(c_int) :: npp(nsp), lpp(FB_MAXSH, nsp)
        real(c_double) :: rpp(nsp), clv(FB_MAXSH, nsp)
        if (all(species(:)%nssh_PP .eq. 0)) return
        lpp = 0
        clv = 0.0d0
        do ispecies = 1, nsp
          npp(ispecies) = species(ispecies)%nssh_PP
          if (npp(ispecies) .gt. FB_MAXSH) error stop ' fireball_hip: more than 4 PP shells '
          rpp(ispecies) = species(ispecies)%rcutoff_PP
          do issh = 1, species(ispecies)%nssh_PP
            lpp(issh, ispecies) = species(ispecies)%shell_PP(issh)%lssh
            clv(issh, ispecies) = species(ispecies)%shell_PP(issh)%cl
          end do
        end do
        call fb_check (fb_set_species_pp(fbh, npp, lpp, rpp, clv), fbh, 'fb_set_species_pp')
        end subroutine set_species_pp

! The library's pseudopotential lists for the current geometry, checked entry
! by entry against the caller's driver_neighbors_PP (s%neighbors_PP, _PPx,
! _PPp with their maps, points and selfs, and the PP common neighbours).
        subroutine fb_session_pp (s)
        type(T_structure), target :: s
        integer(c_int64_t) :: npp, nppp, ntrip, sz(6)
        integer(c_int), allocatable, target :: n(:), self(:), a(:, :)
        integer iatom, p, m, which
        logical ok
        if (have_pp_lists) return
        call fb_check (fb_build_lists_pp(fbh, npp, nppp, ntrip), fbh, 'fb_build_lists_pp')
        call fb_check (fb_get_pp_list_sizes(fbh, 0, sz), fbh, 'fb_get_pp_list_sizes')
        allocate (n(s%natoms), self(s%natoms))
        ok = .true.
        do which = 0, 3
          allocate (a(max(sz(which+1), 1_8), 5))
          call fb_check (fb_get_pp_lists(fbh, 0, which, c_loc(n), c_loc(a(1, 1)), c_loc(a(1, 2)),   &
     &      c_loc(a(1, 3)), c_loc(a(1, 4)), c_loc(a(1, 5)), c_loc(self)), fbh, 'fb_get_pp_lists')
          p = 0
          do iatom = 1, s%natoms
            select case (which)
            case (0)
              m = s%neighbors_PP(iatom)%neighn
              if (m .ne. n(iatom) .or. self(iatom) .ne. s%neighbors_PP_self(iatom)) ok = .false.
              if (.not. ok) exit
              if (any(s%neighbors_PP(iatom)%neigh_j(1:m) .ne. a(p+1:p+m, 1))) ok = .false.
              if (any(s%neighbors_PP(iatom)%neigh_b(1:m) .ne. a(p+1:p+m, 2))) ok = .false.
              if (any(s%neighbors_PP(iatom)%map(1:m) .ne. a(p+1:p+m, 3))) ok = .false.
            case (1)
              m = s%neighbors_PPx(iatom)%neighn
              if (m .ne. n(iatom) .or. self(iatom) .ne. s%neighbors_PPx_self(iatom)) ok = .false.
              if (.not. ok) exit
              if (any(s%neighbors_PPx(iatom)%neigh_j(1:m) .ne. a(p+1:p+m, 1))) ok = .false.
              if (any(s%neighbors_PPx(iatom)%neigh_b(1:m) .ne. a(p+1:p+m, 2))) ok = .false.
              if (any(s%neighbors_PPx(iatom)%map(1:m) .ne. a(p+1:p+m, 3))) ok = .false.
              if (any(s%neighbors_PPx(iatom)%point(1:m) .ne. a(p+1:p+m, 4))) ok = .false.
            case (2)
              m = s%neighbors_PPp(iatom)%neighn
              if (m .ne. n(iatom) .or. self(iatom) .ne. s%neighbors_PPp_self(iatom)) ok = .false.
              if (.not. ok) exit
              if (any(s%neighbors_PPp(iatom)%neigh_j(1:m) .ne. a(p+1:p+m, 1))) ok = .false.
              if (any(s%neighbors_PPp(iatom)%neigh_b(1:m) .ne. a(p+1:p+m, 2))) ok = .false.
            case (3)
              m = s%neighbors_PP(iatom)%ncommon
              if (m .ne. n(iatom)) ok = .false.
              if (.not. ok) exit
              if (m .gt. 0) then
                if (any(s%neighbors_PP(iatom)%iatom_common_j(1:m) .ne. a(p+1:p+m, 1))) ok = .false.
                if (any(s%neighbors_PP(iatom)%iatom_common_b(1:m) .ne. a(p+1:p+m, 2))) ok = .false.
                if (any(s%neighbors_PP(iatom)%jatom_common_j(1:m) .ne. a(p+1:p+m, 3))) ok = .false.
                if (any(s%neighbors_PP(iatom)%jatom_common_b(1:m) .ne. a(p+1:p+m, 4))) ok = .false.
                if (any(s%neighbors_PP(iatom)%neigh_common(1:m) .ne. a(p+1:p+m, 5))) ok = .false.
              end if
            end select
            p = p + m
          end do
          deallocate (a)
          if (.not. ok) exit
        end do
        deallocate (n, self)
        if (.not. ok) then
          write (*,*) ' fireball_hip: the GPU PP lists differ from s%neighbors_PP/_PPx/_PPp '
          error stop ' fireball_hip failure '
        end if
        have_pp_lists = .true.
        end subroutine fb_session_pp

! Fill the blocks of s%svnl (which = 0, per PP pair, norb_mu x norb_PP_nu) or
! s%vnl (which = 1, per PPp pair, norb_mu x norb_nu) from the library,
! allocating them when asked.
        subroutine fb_session_ppstore (s, m, which, alloc)
        type(T_structure), target :: s
        type(T_assemble_neighbors), pointer :: m(:)
        integer, intent(in) :: which
        logical, intent(in) :: alloc
        integer iatom, ineigh, jatom, nmu, nnu, nb, nn
        integer(c_int64_t) :: sz(6), off
        real(c_double), allocatable :: w(:)
        type(T_assemble_block), pointer :: pb
        call fb_check (fb_get_pp_list_sizes(fbh, 0, sz), fbh, 'fb_get_pp_list_sizes')
        allocate (w(max(sz(5 + which), 1_8)))
        call fb_check (fb_get_pp_blocks(fbh, 0, which, w), fbh, 'fb_get_pp_blocks')
        off = 0
        do iatom = 1, s%natoms
          nmu = species(s%atom(iatom)%imass)%norb_max
          if (which .eq. 0) then
            nn = s%neighbors_PP(iatom)%neighn
          else
            nn = s%neighbors_PPp(iatom)%neighn
          end if
          if (alloc) allocate (m(iatom)%neighbors(nn))
          do ineigh = 1, nn
            if (which .eq. 0) then
              jatom = s%neighbors_PP(iatom)%neigh_j(ineigh)
              nnu = species(s%atom(jatom)%imass)%norb_PP_max
            else
              jatom = s%neighbors_PPp(iatom)%neigh_j(ineigh)
              nnu = species(s%atom(jatom)%imass)%norb_max
            end if
            nb = nmu*nnu
            pb => m(iatom)%neighbors(ineigh)
            if (alloc) allocate (pb%block(nmu, nnu))
            pb%block = reshape(w(off+1:off+nb), [nmu, nnu])
            off = off + nb
          end do
        end do
        deallocate (w)
        end subroutine fb_session_ppstore

! Hand the current shell charges to the library (Vna only depends on them).
        subroutine fb_session_charges (s)
        type(T_structure), target :: s
        real(c_double), allocatable :: dq(:, :)
        allocate (dq(FB_MAXSH, s%natoms))
        call shell_charges (s, dq)
        call fb_check (fb_set_charges(fbh, 0, dq), fbh, 'fb_set_charges')
        deallocate (dq)
        end subroutine fb_session_charges

! Hand the input shell charges s%atom(:)%shell(:)%Qin (McWEDA densities) to the library.
        subroutine fb_session_qin (s)
        type(T_structure), target :: s
        real(c_double), allocatable :: q(:, :)
        integer iatom, issh
        allocate (q(FB_MAXSH, s%natoms))
        q = 0.0d0
        do iatom = 1, s%natoms
          do issh = 1, min(species(s%atom(iatom)%imass)%nssh, FB_MAXSH)
            q(issh, iatom) = s%atom(iatom)%shell(issh)%Qin
          end do
        end do
        call fb_check (fb_set_qin(fbh, 0, q), fbh, 'fb_set_qin')
        deallocate (q)
        end subroutine fb_session_qin

! Fill the shell blocks (nssh_mu x nssh_nu) of one T_assemble_neighbors array
! from library matrix imat (allocating them when asked).
        subroutine fb_session_wstore (s, m, imat, alloc)
        type(T_structure), target :: s
        type(T_assemble_neighbors), pointer :: m(:)
        integer(c_int), intent(in) :: imat
        logical, intent(in) :: alloc
        integer iatom, ineigh, jatom, nmu, nnu, nb
        integer(c_int64_t) :: wlen, off
        real(c_double), allocatable :: w(:)
        type(T_assemble_block), pointer :: pb
        call fb_check (fb_get_shell_block_len(fbh, 0, wlen), fbh, 'fb_get_shell_block_len')
        allocate (w(max(wlen, 1_8)))
        call fb_check (fb_get_blocks(fbh, 0, imat, w), fbh, 'fb_get_blocks')
        off = 0
        do iatom = 1, s%natoms
          nmu = species(s%atom(iatom)%imass)%nssh
          if (alloc) allocate (m(iatom)%neighbors(s%neighbors(iatom)%neighn))
          do ineigh = 1, s%neighbors(iatom)%neighn
            jatom = s%neighbors(iatom)%neigh_j(ineigh)
            nnu = species(s%atom(jatom)%imass)%nssh
            nb = nmu*nnu
            pb => m(iatom)%neighbors(ineigh)
            if (alloc) allocate (pb%block(nmu, nnu))
            pb%block = reshape(w(off+1:off+nb), [nmu, nnu])
            off = off + nb
          end do
        end do
        deallocate (w)
        end subroutine fb_session_wstore

        subroutine shell_charges (s, dq)
        type(T_structure), target :: s
        real(c_double), intent(out) :: dq(FB_MAXSH, s%natoms)
        integer iatom, issh
        dq = 0.0d0
        do iatom = 1, s%natoms
          if (.not. associated(s%atom(iatom)%shell)) cycle
          do issh = 1, min(size(s%atom(iatom)%shell), FB_MAXSH)
            dq(issh, iatom) = s%atom(iatom)%shell(issh)%dQ
          end do
        end do
        end subroutine shell_charges

! The library's lists against the caller's: neighbors (neighn, neigh_j,
! neigh_b) and common neighbors (ncommon and the five per-triplet arrays).
        subroutine check_lists (s)
        type(T_structure), target :: s
        integer(c_int), allocatable :: nn(:), nj(:), nb(:), nback(:), nself(:)
        integer(c_int), allocatable :: nc(:), cij(:), cib(:), cjj(:), cjb(:), cm(:)
        integer iatom, p, n
        logical ok

        allocate (nn(s%natoms), nself(s%natoms), nc(s%natoms))
        allocate (nj(max(npairs_all, 1_8)), nb(max(npairs_all, 1_8)), nback(max(npairs_all, 1_8)))
        allocate (cij(max(ntrip_all, 1_8)), cib(max(ntrip_all, 1_8)), cjj(max(ntrip_all, 1_8)))
        allocate (cjb(max(ntrip_all, 1_8)), cm(max(ntrip_all, 1_8)))
        call fb_check (fb_get_neighbors(fbh, 0, nn, nj, nb, nback, nself), fbh, 'fb_get_neighbors')
        call fb_check (fb_get_common(fbh, 0, nc, cij, cib, cjj, cjb, cm), fbh, 'fb_get_common')
        ok = .true.
        p = 0
        do iatom = 1, s%natoms
          n = s%neighbors(iatom)%neighn
          if (n .ne. nn(iatom)) then
            ok = .false.
            exit
          end if
          if (any(s%neighbors(iatom)%neigh_j(1:n) .ne. nj(p+1:p+n))) ok = .false.
          if (any(s%neighbors(iatom)%neigh_b(1:n) .ne. nb(p+1:p+n))) ok = .false.
          p = p + n
        end do
        if (ok) then
          p = 0
          do iatom = 1, s%natoms
            n = s%neighbors(iatom)%ncommon
            if (n .ne. nc(iatom)) then
              ok = .false.
              exit
            end if
            if (n .eq. 0) cycle
            if (any(s%neighbors(iatom)%iatom_common_j(1:n) .ne. cij(p+1:p+n))) ok = .false.
            if (any(s%neighbors(iatom)%iatom_common_b(1:n) .ne. cib(p+1:p+n))) ok = .false.
            if (any(s%neighbors(iatom)%jatom_common_j(1:n) .ne. cjj(p+1:p+n))) ok = .false.
            if (any(s%neighbors(iatom)%jatom_common_b(1:n) .ne. cjb(p+1:p+n))) ok = .false.
            if (any(s%neighbors(iatom)%neigh_common(1:n) .ne. cm(p+1:p+n))) ok = .false.
            p = p + n
          end do
        end if
        deallocate (nn, nj, nb, nback, nself, nc, cij, cib, cjj, cjb, cm)
        if (.not. ok) then
          write (*,*) ' fireball_hip: the GPU neighbor lists differ from s%neighbors'
          error stop ' fireball_hip failure '
        end if
        end subroutine check_lists

! Fetch one assembled matrix (pair order of s%neighbors, each block column-major
! norb_mu x norb_nu) into the session buffer.
        subroutine fb_session_fetch (imat)
        integer(c_int), intent(in) :: imat
        integer(c_int64_t) :: blen
        call fb_check (fb_get_block_len(fbh, 0, blen), fbh, 'fb_get_block_len')
        if (allocated(fb_buf)) then
          if (size(fb_buf, kind=8) .lt. blen) deallocate (fb_buf)
        end if
        if (.not. allocated(fb_buf)) allocate (fb_buf(max(blen, 1_8)))
        call fb_check (fb_get_blocks(fbh, 0, imat, fb_buf), fbh, 'fb_get_blocks')
        end subroutine fb_session_fetch

! Allocate (when asked) and fill the blocks of one T_assemble_neighbors array
! from the session buffer; which = 0 -> %block, 1 -> %blocko.
        subroutine fb_session_store (s, m, which, alloc)
        type(T_structure), target :: s
        type(T_assemble_neighbors), pointer :: m(:)
        integer, intent(in) :: which
        logical, intent(in) :: alloc
        integer iatom, ineigh, jatom, nmu, nnu, nb
        integer(c_int64_t) :: off
        type(T_assemble_block), pointer :: pb

        off = 0
        do iatom = 1, s%natoms
          nmu = species(s%atom(iatom)%imass)%norb_max
          if (alloc .and. which .eq. 0) allocate (m(iatom)%neighbors(s%neighbors(iatom)%neighn))
          do ineigh = 1, s%neighbors(iatom)%neighn
            jatom = s%neighbors(iatom)%neigh_j(ineigh)
            nnu = species(s%atom(jatom)%imass)%norb_max
            nb = nmu*nnu
            pb => m(iatom)%neighbors(ineigh)
            if (which .eq. 0) then
              if (alloc) allocate (pb%block(nmu, nnu))
              pb%block = reshape(fb_buf(off+1:off+nb), [nmu, nnu])
            else
              if (alloc) allocate (pb%blocko(nmu, nnu))
              pb%blocko = reshape(fb_buf(off+1:off+nb), [nmu, nnu])
            end if
            off = off + nb
          end do
        end do
        end subroutine fb_session_store

! Remember the Vna blocks just stored (the session buffer) as the caller's.
        subroutine fb_session_keep_vna ()
        if (allocated(vna_last)) deallocate (vna_last)
        allocate (vna_last(size(fb_buf, kind=8)))
        vna_last = fb_buf
        end subroutine fb_session_keep_vna

! Add the change of the library's Vna blocks (session buffer) since they were
! last handed over into the caller's s%vna(:)%neighbors(:)%block.
        subroutine fb_session_add_vna (s)
        type(T_structure), target :: s
        integer iatom, ineigh, jatom, nmu, nnu, nb
        integer(c_int64_t) :: off
        type(T_assemble_block), pointer :: pb
        if (.not. allocated(vna_last)) error stop ' fireball_hip: assemble_vna_3c before assemble_vna_2c '
        off = 0
        do iatom = 1, s%natoms
          nmu = species(s%atom(iatom)%imass)%norb_max
          do ineigh = 1, s%neighbors(iatom)%neighn
            jatom = s%neighbors(iatom)%neigh_j(ineigh)
            nnu = species(s%atom(jatom)%imass)%norb_max
            nb = nmu*nnu
            pb => s%vna(iatom)%neighbors(ineigh)
            pb%block = pb%block + reshape(fb_buf(off+1:off+nb) - vna_last(off+1:off+nb), [nmu, nnu])
            off = off + nb
          end do
        end do
        vna_last = fb_buf
        end subroutine fb_session_add_vna

! Hand the caller's density matrix (s%denmat, the blocks the force assemblers
! contract with) to the library, in the pair order of s%neighbors.
        subroutine fb_session_denmat (s)
        type(T_structure), target :: s
        integer iatom, ineigh, nb
        integer(c_int64_t) :: blen, off
        real(c_double), allocatable :: rho(:)
        call fb_check (fb_get_block_len(fbh, 0, blen), fbh, 'fb_get_block_len')
        allocate (rho(max(blen, 1_8)))
        off = 0
        do iatom = 1, s%natoms
          do ineigh = 1, s%neighbors(iatom)%neighn
            nb = size(s%denmat(iatom)%neighbors(ineigh)%block)
            rho(off+1:off+nb) = reshape(s%denmat(iatom)%neighbors(ineigh)%block, [nb])
            off = off + nb
          end do
        end do
        if (off .ne. blen) error stop ' fireball_hip: s%denmat does not match the block layout '
        call fb_check (fb_set_denmat(fbh, 0, rho), fbh, 'fb_set_denmat')
        deallocate (rho)
        end subroutine fb_session_denmat

! Allocate and fill Dblock(3, norb_mu, norb_nu) of one T_assemble_neighbors
! array from the library's Dblocks of matrix imat.
        subroutine fb_session_dstore (s, m, imat)
        type(T_structure), target :: s
        type(T_assemble_neighbors), pointer :: m(:)
        integer(c_int), intent(in) :: imat
        integer iatom, ineigh, jatom, nmu, nnu, nb
        integer(c_int64_t) :: blen, off
        real(c_double), allocatable :: d(:)
        type(T_assemble_block), pointer :: pb
        call fb_check (fb_get_block_len(fbh, 0, blen), fbh, 'fb_get_block_len')
        allocate (d(max(3*blen, 1_8)))
        call fb_check (fb_get_dblocks(fbh, 0, imat, d), fbh, 'fb_get_dblocks')
        off = 0
        do iatom = 1, s%natoms
          nmu = species(s%atom(iatom)%imass)%norb_max
          do ineigh = 1, s%neighbors(iatom)%neighn
            jatom = s%neighbors(iatom)%neigh_j(ineigh)
            nnu = species(s%atom(jatom)%imass)%norb_max
            nb = 3*nmu*nnu
            pb => m(iatom)%neighbors(ineigh)
            allocate (pb%Dblock(3, nmu, nnu))
            pb%Dblock = reshape(d(off+1:off+nb), [3, nmu, nnu])
            off = off + nb
          end do
        end do
        deallocate (d)
        end subroutine fb_session_dstore

        end module M_fb_hip_session
