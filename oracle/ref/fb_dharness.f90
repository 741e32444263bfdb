! Force harness of this project (not reference code).
!
! Same set-up as fb_harness.f90 (tables, cells, neighbour lists, assemble_S,
! assemble_T, assemble_dipole_z, assemble_vna_2c, assemble_vna_3c), then the
! force assemblers of k.DASSEMBLERS: Dassemble_S, Dassemble_T,
! Dassemble_dipole_z, Dassemble_vna_2c (module M_Dassemble_2c) and
! Dassemble_vna_3c (module M_Dassemble_3c), in the order the reference's MD
! driver calls them (n.NAMD/M_nonadiabatic_mdet.f90:2065-2101), with a density
! matrix read from a file in place of the SCF solution.
!
! The same source links against either the reference's M_Dassemble_2c /
! M_Dassemble_3c objects or the drop-in modules of this project.
!
! Usage:  fb_dharness <input> <denmat> <output> [nrep]
!   input  : written by fireball_hip.fdata_io.write_harness_input
!   denmat : stream of binary64 density blocks, pair order (iatom major,
!            ineigh minor), each column-major norb_mu x norb_nu
!   output : stream-unformatted dump of Dblocks, force pieces and timings
!   nrep   : repeat the Dassemble_* sequence nrep times for timing (default 1)
        program fb_dharness
        use M_precision
        use M_species
        use M_configuraciones
        use M_neighbors
        use M_Fdata_2c
        use M_Fdata_3c
        use M_assemble_2c
        use M_assemble_3c
        use M_Dassemble_2c
        use M_Dassemble_3c
        implicit none

        type(T_structure), target :: st
        integer ispecies, issh, nssh, iatom, ineigh, irep, nrep, nargs
        integer in1, in2, ntot_pairs, ntot_common, nmu, nnu
        integer(kind=8) :: c0, c1, crate
        real t_S, t_T, t_dip, t_vna2c, t_vna3c
        real(kind=8) dqscale
        character(len=512) :: infile, rhofile, outfile, arg
        character(len=128) :: floc

        nargs = command_argument_count()
        call get_command_argument(1, infile)
        call get_command_argument(2, rhofile)
        call get_command_argument(3, outfile)
        nrep = 1
        if (nargs .ge. 4) then
          call get_command_argument(4, arg)
          read (arg, *) nrep
        end if
        ! optional: scale every shell charge after the H+S assembly, as a
        ! charge mixer does between the last assemble_vna and the force calls
        dqscale = 1.0d0
        if (nargs .ge. 5) then
          call get_command_argument(5, arg)
          read (arg, *) dqscale
        end if

        open (unit = ilogfile, file = 'fb_dharness.log', status = 'unknown')
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
        call assemble_S (st)
        call assemble_T (st)
        call assemble_dipole_z (st)
        call assemble_vna_2c (st)
        call assemble_vna_3c (st)
        if (dqscale .ne. 1.0d0) then
          do iatom = 1, st%natoms
            do issh = 1, species(st%atom(iatom)%imass)%nssh
              st%atom(iatom)%shell(issh)%dQ = st%atom(iatom)%shell(issh)%dQ*dqscale
            end do
          end do
        end if

        ! density matrix blocks, same layout as the H/S blocks
        allocate (st%denmat(st%natoms))
        open (unit = 32, file = trim(rhofile), access = 'stream',               &
     &        form = 'unformatted', status = 'old')
        do iatom = 1, st%natoms
          in1 = st%atom(iatom)%imass
          nmu = species(in1)%norb_max
          allocate (st%denmat(iatom)%neighbors(st%neighbors(iatom)%neighn))
          do ineigh = 1, st%neighbors(iatom)%neighn
            in2 = st%atom(st%neighbors(iatom)%neigh_j(ineigh))%imass
            nnu = species(in2)%norb_max
            allocate (st%denmat(iatom)%neighbors(ineigh)%block(nmu, nnu))
            read (32) st%denmat(iatom)%neighbors(ineigh)%block
          end do
        end do
        close (32)
        allocate (st%forces(st%natoms))

        call system_clock (count_rate = crate)
        t_S = 0.0d0; t_T = 0.0d0; t_dip = 0.0d0; t_vna2c = 0.0d0; t_vna3c = 0.0d0
        do irep = 1, nrep
          if (irep .gt. 1) call destroy_Dassemble_2c (st)
          do iatom = 1, st%natoms
            st%forces(iatom)%f3naa = 0.0d0
            st%forces(iatom)%f3nab = 0.0d0
            st%forces(iatom)%f3nac = 0.0d0
          end do
          call system_clock (c0)
          call Dassemble_S (st)
          call system_clock (c1)
          t_S = t_S + real(c1 - c0)/real(crate)
          call system_clock (c0)
          call Dassemble_T (st)
          call system_clock (c1)
          t_T = t_T + real(c1 - c0)/real(crate)
          call system_clock (c0)
          call Dassemble_dipole_z (st)
          call system_clock (c1)
          t_dip = t_dip + real(c1 - c0)/real(crate)
          call system_clock (c0)
          call Dassemble_vna_2c (st)
          call system_clock (c1)
          t_vna2c = t_vna2c + real(c1 - c0)/real(crate)
          call system_clock (c0)
          call Dassemble_vna_3c (st)
          call system_clock (c1)
          t_vna3c = t_vna3c + real(c1 - c0)/real(crate)
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
        write (31) t_S, t_T, t_dip, t_vna2c, t_vna3c
        ! Dblock(3, norb_mu, norb_nu) of S, T, dipole_z per (iatom, ineigh)
        do iatom = 1, st%natoms
          do ineigh = 1, st%neighbors(iatom)%neighn
            write (31) st%overlap(iatom)%neighbors(ineigh)%Dblock
            write (31) st%kinetic(iatom)%neighbors(ineigh)%Dblock
            write (31) st%dipole_z(iatom)%neighbors(ineigh)%Dblock
          end do
        end do
        do iatom = 1, st%natoms
          write (31) st%forces(iatom)%vna_atom(:, 1:st%neighbors(iatom)%neighn)
          write (31) st%forces(iatom)%vna_ontop(:, 1:st%neighbors(iatom)%neighn)
        end do
        do iatom = 1, st%natoms
          write (31) st%forces(iatom)%f3naa, st%forces(iatom)%f3nab,          &
     &               st%forces(iatom)%f3nac
        end do
        close (31)
        write (*, '(A, 5F14.6)') ' times (st): DS DT Ddip Dvna2c Dvna3c ',    &
     &    t_S, t_T, t_dip, t_vna2c, t_vna3c
        end program fb_dharness
