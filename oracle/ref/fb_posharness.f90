! Test harness of this project (not reference code).
!
! Runs the reference's read_positions (c.SYSTEM/M_configuraciones.f90:604-797)
! on a structure file and dumps what it leaves in T_structure: natoms,
! icluster, the lattice, ishiftO (whether the origin shift was applied), the
! species index and position of every atom, norbitals and iblock_slot.  The
! species come from a small input file instead of info.dat.
!
! Usage:  fb_posharness <species input> <structure file> <output>
!   species input: nspecies, then per species "nZ xmass norb_max"
        program fb_posharness
        use M_precision
        use M_species
        use M_configuraciones
        implicit none

        type(T_structure), target :: st
        integer ispecies, iatom
        character(len=512) :: spfile, basis, outfile

        call get_command_argument(1, spfile)
        call get_command_argument(2, basis)
        call get_command_argument(3, outfile)

        open (unit = ilogfile, file = 'fb_posharness.log', status = 'unknown')
        open (unit = 30, file = trim(spfile), status = 'old')
        read (30, *) nspecies
        allocate (species(nspecies))
        do ispecies = 1, nspecies
          read (30, *) species(ispecies)%nZ, species(ispecies)%xmass,        &
     &                 species(ispecies)%norb_max
          species(ispecies)%symbol = 'Xx'
        end do
        close (30)

        st%logfile = ilogfile
        st%inpfile = 31
        st%basisfile = basis
        call read_positions (st)

        open (unit = 32, file = trim(outfile), access = 'stream',              &
     &        form = 'unformatted', status = 'replace')
        write (32) st%natoms, st%icluster, ishiftO, st%norbitals
        do iatom = 1, 3
          write (32) st%lattice(iatom)%a
        end do
        do iatom = 1, st%natoms
          write (32) st%atom(iatom)%imass, st%iblock_slot(iatom)
          write (32) st%atom(iatom)%ratom
        end do
        close (32)
        end program fb_posharness
