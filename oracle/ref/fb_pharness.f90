! Pseudopotential harness of this project (not reference code).
!
! Same set-up as fb_harness.f90, with the pseudopotential shells of every
! species read from a second file (nssh_PP, lssh_PP, rcutoff_PP; the
! Kleinman-Bylander cl values come from clPP.ZZ.dat through read_Fdata_2c),
! then driver_neighbors_PP (c.SYSTEM/M_neighbors_PP.f90), assemble_svnl and
! assemble_vnl_2c (module M_assemble_PP_2c) and assemble_vnl_3c (module
! M_assemble_PP_3c).
!
! The same source links against either the reference's modules or the
! drop-in modules of this project.
!
! Usage:  fb_pharness <input> <ppinput> <output> [nrep]
!   input   : written by fireball_hip.fdata_io.write_harness_input
!   ppinput : written by fireball_hip.fdata_io.write_pp_input
!   output  : stream-unformatted dump of the PP lists, sVNL and vnl blocks, timings
!   nrep    : repeat the assemblers nrep times for timing (default 1)
        program fb_pharness
        use M_precision
        use M_species
        use M_configuraciones
        use M_neighbors
        use M_neighbors_PP
        use M_Fdata_2c
        use M_Fdata_3c
        use M_assemble_PP_2c
        use M_assemble_PP_3c
        implicit none

        type(T_structure), target :: st
        integer ispecies, issh, nssh, iatom, ineigh, irep, nrep, nargs, lqn, n
        integer(kind=8) :: c0, c1, crate
        real t_sv, t_v2, t_v3
        character(len=512) :: infile, ppfile, outfile, arg
        character(len=128) :: floc

        nargs = command_argument_count()
        call get_command_argument(1, infile)
        call get_command_argument(2, ppfile)
        call get_command_argument(3, outfile)
        nrep = 1
        if (nargs .ge. 4) then
          call get_command_argument(4, arg)
          read (arg, *) nrep
        end if

        open (unit = ilogfile, file = 'fb_pharness.log', status = 'unknown')
        open (unit = 30, file = trim(infile), status = 'old')
        read (30, *) floc
        Fdata_location = floc
        read (30, *) nspecies
        allocate (species(nspecies))
        have_dorbitals = .false.
        do ispecies = 1, nspecies
          read (30, *) species(ispecies)%nZ, nssh
          species(ispecies)%nssh = nssh
          species(ispecies)%nssh_PP = 0
          allocate (species(ispecies)%shell(nssh))
          allocate (species(ispecies)%shell_PP(0))
          read (30, *) (species(ispecies)%shell(issh)%lssh, issh = 1, nssh)
          read (30, *) (species(ispecies)%shell(issh)%rcutoff, issh = 1, nssh)
          species(ispecies)%norb_max = 0
          do issh = 1, nssh
            species(ispecies)%shell(issh)%rcutoffA =                          &
     &        species(ispecies)%shell(issh)%rcutoff*0.529177d0
            species(ispecies)%norb_max = species(ispecies)%norb_max          &
     &        + 2*species(ispecies)%shell(issh)%lssh + 1
            if (species(ispecies)%shell(issh)%lssh .eq. 2) have_dorbitals = .true.
          end do
        end do

        read (30, *) st%natoms, st%icluster
        do issh = 1, 3
          read (30, *) st%lattice(issh)%a
        end do
        allocate (st%atom(st%natoms))
        do iatom = 1, st%natoms
          read (30, *) st%atom(iatom)%imass, st%atom(iatom)%ratom
          st%atom(iatom)%species => species(st%atom(iatom)%imass)
        end do
        do iatom = 1, st%natoms
          nssh = species(st%atom(iatom)%imass)%nssh
          allocate (st%atom(iatom)%shell(nssh))
          read (30, *) (st%atom(iatom)%shell(issh)%dQ, issh = 1, nssh)
        end do
        close (30)
        st%logfile = ilogfile
        st%basisfile = 'harness.inp'
        ifix_neighbors = 0
        iwriteout_neighbors = 0

        ! pseudopotential shells
        open (unit = 33, file = trim(ppfile), status = 'old')
        do ispecies = 1, nspecies
          read (33, *) nssh
          deallocate (species(ispecies)%shell_PP)
          allocate (species(ispecies)%shell_PP(nssh))
          species(ispecies)%nssh_PP = nssh
          read (33, *) (species(ispecies)%shell_PP(issh)%lssh, issh = 1, nssh)
          read (33, *) species(ispecies)%rcutoff_PP
          species(ispecies)%norb_PP_max = 0
          do issh = 1, nssh
            lqn = species(ispecies)%shell_PP(issh)%lssh
            species(ispecies)%norb_PP_max = species(ispecies)%norb_PP_max + 2*lqn + 1
          end do
        end do
        close (33)

        call read_Fdata_2c ()
        call read_Fdata_3c ()
        call make_cells (st)
        call driver_neighbors (st)
        call driver_neighbors_PP (st)

        call system_clock (count_rate = crate)
        t_sv = 0.0d0; t_v2 = 0.0d0; t_v3 = 0.0d0
        do irep = 1, nrep
          if (irep .gt. 1) call destroy_assemble_PP_2c (st)
          call system_clock (c0)
          call assemble_svnl (st)
          call system_clock (c1)
          t_sv = t_sv + real(c1 - c0)/real(crate)
          call system_clock (c0)
          call assemble_vnl_2c (st)
          call system_clock (c1)
          t_v2 = t_v2 + real(c1 - c0)/real(crate)
          call system_clock (c0)
          call assemble_vnl_3c (st)
          call system_clock (c1)
          t_v3 = t_v3 + real(c1 - c0)/real(crate)
        end do

        open (unit = 31, file = trim(outfile), access = 'stream',              &
     &        form = 'unformatted', status = 'replace')
        write (31) st%natoms, nrep
        write (31) t_sv, t_v2, t_v3
        ! neighbors_PP: neighn, per atom neigh_j, neigh_b, map; neighbors_PP_self
        write (31) (st%neighbors_PP(iatom)%neighn, iatom = 1, st%natoms)
        do iatom = 1, st%natoms
          n = st%neighbors_PP(iatom)%neighn
          write (31) st%neighbors_PP(iatom)%neigh_j(1:n), st%neighbors_PP(iatom)%neigh_b(1:n),   &
     &               st%neighbors_PP(iatom)%map(1:n)
        end do
        write (31) st%neighbors_PP_self(1:st%natoms)
        ! neighbors_PPx: neighn, per atom neigh_j, neigh_b, map, point; self
        write (31) (st%neighbors_PPx(iatom)%neighn, iatom = 1, st%natoms)
        do iatom = 1, st%natoms
          n = st%neighbors_PPx(iatom)%neighn
          write (31) st%neighbors_PPx(iatom)%neigh_j(1:n), st%neighbors_PPx(iatom)%neigh_b(1:n), &
     &               st%neighbors_PPx(iatom)%map(1:n), st%neighbors_PPx(iatom)%point(1:n)
        end do
        write (31) st%neighbors_PPx_self(1:st%natoms)
        ! neighbors_PPp: neighn, per atom neigh_j, neigh_b; self
        write (31) (st%neighbors_PPp(iatom)%neighn, iatom = 1, st%natoms)
        do iatom = 1, st%natoms
          n = st%neighbors_PPp(iatom)%neighn
          write (31) st%neighbors_PPp(iatom)%neigh_j(1:n), st%neighbors_PPp(iatom)%neigh_b(1:n)
        end do
        write (31) st%neighbors_PPp_self(1:st%natoms)
        ! PP common neighbours: ncommon, per atom the five arrays
        write (31) (st%neighbors_PP(iatom)%ncommon, iatom = 1, st%natoms)
        do iatom = 1, st%natoms
          n = st%neighbors_PP(iatom)%ncommon
          write (31) st%neighbors_PP(iatom)%iatom_common_j(1:n),                      &
     &               st%neighbors_PP(iatom)%iatom_common_b(1:n),                      &
     &               st%neighbors_PP(iatom)%jatom_common_j(1:n),                      &
     &               st%neighbors_PP(iatom)%jatom_common_b(1:n),                      &
     &               st%neighbors_PP(iatom)%neigh_common(1:n)
        end do
        ! sVNL per PP pair, vnl per PPp pair
        do iatom = 1, st%natoms
          do ineigh = 1, st%neighbors_PP(iatom)%neighn
            write (31) st%svnl(iatom)%neighbors(ineigh)%block
          end do
        end do
        do iatom = 1, st%natoms
          do ineigh = 1, st%neighbors_PPp(iatom)%neighn
            write (31) st%vnl(iatom)%neighbors(ineigh)%block
          end do
        end do
        close (31)
        write (*, '(A, 3F14.6)') ' times (s): svnl vnl_2c vnl_3c ', t_sv, t_v2, t_v3
        end program fb_pharness
