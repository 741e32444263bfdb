! Test harness of this project (not reference code).
!
! Drives the reference's own modules through the 2c/3c assembly path:
! read_Fdata_2c / read_Fdata_3c (e.FDATA), make_cells (c.SYSTEM/make_cells.f90),
! driver_neighbors (c.SYSTEM/M_neighbors.f90:93-141), then assemble_S,
! assemble_T, assemble_dipole_z, assemble_vna_2c (module M_assemble_2c) and
! assemble_vna_3c (module M_assemble_3c).  The reference tree ships no main
! program, so this harness stands in the place of the missing driver; the
! species and atoms are filled directly instead of through info.dat / the .inp
! reader.
!
! The same source links against either the reference's M_assemble_2c /
! M_assemble_3c objects (golden vectors, CPU baseline) or the drop-in modules
! of this project that call the GPU library; the output files are then
! comparable byte for byte in layout.
!
! Usage:  fb_harness <input> <output> [nrep [vnabump]]
!   input  : written by fireball_hip.fdata_io.write_harness_input
!   output : stream-unformatted dump of lists, blocks and timings
!   nrep   : repeat the assemble_* sequence nrep times for timing (default 1)
!   vnabump: add this value to every element of s%vna(:)%neighbors(:)%block
!            between assemble_vna_2c and assemble_vna_3c (a caller that edits
!            the blocks in between; default 0)
        program fb_harness
        use M_precision
        use M_species
        use M_configuraciones
        use M_neighbors
        use M_Fdata_2c
        use M_Fdata_3c
        use M_assemble_2c
        use M_assemble_3c
        implicit none

        type(T_structure), target :: st
        integer ispecies, issh, nssh, iatom, ineigh, irep, nrep, nargs
        integer in1, in2, ialpha, ntot_pairs, ntot_common
        integer(kind=8) :: c0, c1, crate
        real t_lists, t_S, t_T, t_dip, t_vna2c, t_vna3c
        real t_rep(5)
        real(kind=8) vnabump
        character(len=512) :: infile, outfile, arg
        character(len=128) :: floc

        nargs = command_argument_count()
        call get_command_argument(1, infile)
        call get_command_argument(2, outfile)
        nrep = 1
        if (nargs .ge. 3) then
          call get_command_argument(3, arg)
          read (arg, *) nrep
        end if
        vnabump = 0.0d0
        if (nargs .ge. 4) then
          call get_command_argument(4, arg)
          read (arg, *) vnabump
        end if

        open (unit = ilogfile, file = 'fb_harness.log', status = 'unknown')
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

        call system_clock (count_rate = crate)
        call system_clock (c0)
        call driver_neighbors (st)
        call system_clock (c1)
        t_lists = real(c1 - c0)/real(crate)

        t_S = 0.0d0; t_T = 0.0d0; t_dip = 0.0d0; t_vna2c = 0.0d0; t_vna3c = 0.0d0
        do irep = 1, nrep
          t_rep = (/t_S, t_T, t_dip, t_vna2c, t_vna3c/)
          if (irep .gt. 1) then
            call destroy_assemble_2c (st)
            call destroy_assemble_vna (st)
          end if
          call system_clock (c0)
          call assemble_S (st)
          call system_clock (c1)
          t_S = t_S + real(c1 - c0)/real(crate)
          call system_clock (c0)
          call assemble_T (st)
          call system_clock (c1)
          t_T = t_T + real(c1 - c0)/real(crate)
          call system_clock (c0)
          call assemble_dipole_z (st)
          call system_clock (c1)
          t_dip = t_dip + real(c1 - c0)/real(crate)
          call system_clock (c0)
          call assemble_vna_2c (st)
          call system_clock (c1)
          t_vna2c = t_vna2c + real(c1 - c0)/real(crate)
          if (vnabump .ne. 0.0d0) then
            do iatom = 1, st%natoms
              do ineigh = 1, st%neighbors(iatom)%neighn
                st%vna(iatom)%neighbors(ineigh)%block =                        &
     &            st%vna(iatom)%neighbors(ineigh)%block + vnabump
              end do
            end do
          end if
          call system_clock (c0)
          call assemble_vna_3c (st)
          call system_clock (c1)
          t_vna3c = t_vna3c + real(c1 - c0)/real(crate)
          ! this repetition's times (seconds): S T dip vna2c vna3c
          write (*, '(A, I6, 5F14.6)') ' rep ', irep,                         &
     &      (/t_S, t_T, t_dip, t_vna2c, t_vna3c/) - t_rep
        end do

        ntot_pairs = 0
        ntot_common = 0
        do iatom = 1, st%natoms
          ntot_pairs = ntot_pairs + st%neighbors(iatom)%neighn
          ntot_common = ntot_common + st%neighbors(iatom)%ncommon
        end do

        open (unit = 31, file = trim(outfile), access = 'stream',              &
     &        form = 'unformatted', status = 'replace')
        write (31) st%natoms, mbeta_max, ntot_pairs, ntot_common, nrep
        write (31) t_lists, t_S, t_T, t_dip, t_vna2c, t_vna3c
        do iatom = 1, st%natoms
          write (31) st%neighbors(iatom)%neighn
        end do
        do iatom = 1, st%natoms
          write (31) st%neighbors(iatom)%neigh_j(1:st%neighbors(iatom)%neighn)
          write (31) st%neighbors(iatom)%neigh_b(1:st%neighbors(iatom)%neighn)
          write (31) st%neighbors(iatom)%neigh_back(1:st%neighbors(iatom)%neighn)
        end do
        write (31) st%neigh_self(1:st%natoms)
        do ialpha = 1, st%natoms
          write (31) st%neighbors(ialpha)%ncommon
        end do
        do ialpha = 1, st%natoms
          ineigh = st%neighbors(ialpha)%ncommon
          write (31) st%neighbors(ialpha)%iatom_common_j(1:ineigh)
          write (31) st%neighbors(ialpha)%iatom_common_b(1:ineigh)
          write (31) st%neighbors(ialpha)%jatom_common_j(1:ineigh)
          write (31) st%neighbors(ialpha)%jatom_common_b(1:ineigh)
          write (31) st%neighbors(ialpha)%neigh_common(1:ineigh)
        end do
        ! blocks: S, T, dipole_z, vna%block, vna%blocko per (iatom, ineigh),
        ! each column-major norb_mu x norb_nu
        do iatom = 1, st%natoms
          in1 = st%atom(iatom)%imass
          do ineigh = 1, st%neighbors(iatom)%neighn
            in2 = st%atom(st%neighbors(iatom)%neigh_j(ineigh))%imass
            write (31) st%overlap(iatom)%neighbors(ineigh)%block
            write (31) st%kinetic(iatom)%neighbors(ineigh)%block
            write (31) st%dipole_z(iatom)%neighbors(ineigh)%block
            write (31) st%vna(iatom)%neighbors(ineigh)%block
            write (31) st%vna(iatom)%neighbors(ineigh)%blocko
          end do
        end do
        close (31)
        write (*, '(A, 6F14.6)') ' times (st): lists S T dip vna2c vna3c ',   &
     &    t_lists, t_S, t_T, t_dip, t_vna2c, t_vna3c
        end program fb_harness
