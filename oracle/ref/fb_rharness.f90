! McWEDA density harness of this project (not reference code).
!
! Same set-up as fb_harness.f90 (tables, cells, neighbour lists), input shell
! charges Qin read from a file, then the density assemblers of module
! M_assemble_rho_McWEDA (j.ASSEMBLERS/M_assemble_rho_McWEDA_DOGS.f90):
! assemble_rho_2c, assemble_rho_3c, assemble_rho_weighted_2c and
! assemble_rho_weighted_3c, in the order of assemble_vxc
! (j.ASSEMBLERS/M_assemble_vxc_McWEDA_DOGS.f90:148-154).
!
! The same source links against either the reference's module or the drop-in
! module of this project.
!
! Usage:  fb_rharness <input> <qin> <output> [nrep]
!   input  : written by fireball_hip.fdata_io.write_harness_input
!   qin    : stream of binary64 shell charges, atom order, nssh values per atom
!   output : stream-unformatted dump of the four density arrays and timings
!   nrep   : repeat the sequence nrep times for timing (default 1)
        program fb_rharness
        use M_precision
        use M_species
        use M_configuraciones
        use M_neighbors
        use M_Fdata_2c
        use M_Fdata_3c
        use M_assemble_rho_McWEDA
        implicit none

        type(T_structure), target :: st
        integer ispecies, issh, nssh, iatom, ineigh, irep, nrep, nargs
        integer ntot_pairs, ntot_common
        integer(kind=8) :: c0, c1, crate
        real t_r2, t_r3, t_w2, t_w3
        character(len=512) :: infile, qfile, outfile, arg
        character(len=128) :: floc

        nargs = command_argument_count()
        call get_command_argument(1, infile)
        call get_command_argument(2, qfile)
        call get_command_argument(3, outfile)
        nrep = 1
        if (nargs .ge. 4) then
          call get_command_argument(4, arg)
          read (arg, *) nrep
        end if

        open (unit = ilogfile, file = 'fb_rharness.log', status = 'unknown')
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

        call read_Fdata_2c ()
        call read_Fdata_3c ()
        call make_cells (st)
        call driver_neighbors (st)

        ! input shell charges (s%atom(:)%shell(:)%Qin)
        open (unit = 32, file = trim(qfile), access = 'stream',                 &
     &        form = 'unformatted', status = 'old')
        do iatom = 1, st%natoms
          nssh = species(st%atom(iatom)%imass)%nssh
          read (32) (st%atom(iatom)%shell(issh)%Qin, issh = 1, nssh)
        end do
        close (32)

        call system_clock (count_rate = crate)
        t_r2 = 0.0d0; t_r3 = 0.0d0; t_w2 = 0.0d0; t_w3 = 0.0d0
        do irep = 1, nrep
          if (irep .gt. 1) call destroy_assemble_rho (st)
          call system_clock (c0)
          call assemble_rho_2c (st)
          call system_clock (c1)
          t_r2 = t_r2 + real(c1 - c0)/real(crate)
          call system_clock (c0)
          call assemble_rho_3c (st)
          call system_clock (c1)
          t_r3 = t_r3 + real(c1 - c0)/real(crate)
          call system_clock (c0)
          call assemble_rho_weighted_2c (st)
          call system_clock (c1)
          t_w2 = t_w2 + real(c1 - c0)/real(crate)
          call system_clock (c0)
          call assemble_rho_weighted_3c (st)
          call system_clock (c1)
          t_w3 = t_w3 + real(c1 - c0)/real(crate)
        end do

        ntot_pairs = 0
        ntot_common = 0
        do iatom = 1, st%natoms
          ntot_pairs = ntot_pairs + st%neighbors(iatom)%neighn
          ntot_common = ntot_common + st%neighbors(iatom)%ncommon
        end do

        open (unit = 31, file = trim(outfile), access = 'stream',              &
     &        form = 'unformatted', status = 'replace')
        write (31) st%natoms, ntot_pairs, ntot_common, nrep
        write (31) t_r2, t_r3, t_w2, t_w3
        do iatom = 1, st%natoms
          do ineigh = 1, st%neighbors(iatom)%neighn
            write (31) st%rho_in(iatom)%neighbors(ineigh)%block
            write (31) st%rho_bond(iatom)%neighbors(ineigh)%block
          end do
        end do
        do iatom = 1, st%natoms
          do ineigh = 1, st%neighbors(iatom)%neighn
            write (31) st%rho_in_weighted(iatom)%neighbors(ineigh)%block
            write (31) st%rho_bond_weighted(iatom)%neighbors(ineigh)%block
          end do
        end do
        close (31)
        write (*, '(A, 4F14.6)') ' times (s): rho2c rho3c rhoW2c rhoW3c ',     &
     &    t_r2, t_r3, t_w2, t_w3
        end program fb_rharness
