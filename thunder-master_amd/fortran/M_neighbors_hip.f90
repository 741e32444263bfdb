! Drop-in for the neighbour driver (c.SYSTEM/M_neighbors.f90:93-141): the same
! module name and the routines its callers use -- driver_neighbors,
! writeout_neighbors, destroy_neighbors -- with s%neighbors and s%neigh_self
! filled from the library's lists (fb_build_lists, fb_get_neighbors,
! fb_get_common, through M_fb_hip_session) instead of the reference's search
! over every atom pair of every cell.  The lists are the reference's, entry
! for entry (the library builds them in the reference's order); the arrays
! have the reference's shapes: neigh_b, neigh_j, neigh_back of length neighn,
! the five common-neighbour arrays of length ncommon.
!
! Link this object in place of the caller's M_neighbors.o (its .mod in place of
! M_neighbors.mod).  ifix_neighbors = 1 (a neighbour map read from a file)
! has no counterpart here: the GPU lists follow the positions, so the run
! stops with a message.
        module M_neighbors
        use iso_c_binding
        use fb_hip_c
        use M_configuraciones
        use M_fb_hip_session
        implicit none

        contains

        subroutine driver_neighbors (s)
        type(T_structure), target :: s
        integer(c_int), allocatable :: nn(:), nj(:), nb(:), nback(:), nself(:)
        integer(c_int), allocatable :: nc(:), cij(:), cib(:), cjj(:), cjb(:), cm(:)
        integer iatom, p, n
        logical rebuilt

        write (s%logfile,*) ' Welcome to neighbors - determine neighbor mapping. '
        if (ifix_neighbors .eq. 1) then
          write (*,*) ' fireball_hip: driver_neighbors builds the lists from the positions;'
          write (*,*) ' a fixed neighbour map (ifix_neighbors = 1) is not supported'
          error stop ' fireball_hip failure '
        end if
        rebuilt = fb_session_build (s)

        allocate (nn(s%natoms), nself(s%natoms), nc(s%natoms))
        allocate (nj(max(npairs_all, 1_8)), nb(max(npairs_all, 1_8)), nback(max(npairs_all, 1_8)))
        allocate (cij(max(ntrip_all, 1_8)), cib(max(ntrip_all, 1_8)), cjj(max(ntrip_all, 1_8)))
        allocate (cjb(max(ntrip_all, 1_8)), cm(max(ntrip_all, 1_8)))
        call fb_check (fb_get_neighbors(fbh, 0, nn, nj, nb, nback, nself), fbh, 'fb_get_neighbors')
        call fb_check (fb_get_common(fbh, 0, nc, cij, cib, cjj, cjb, cm), fbh, 'fb_get_common')

        allocate (s%neighbors(s%natoms))
        allocate (s%neigh_self(s%natoms))
        s%neigh_self = nself
        p = 0
        do iatom = 1, s%natoms
          n = nn(iatom)
          s%neighbors(iatom)%neighn = n
          allocate (s%neighbors(iatom)%neigh_b(n))
          allocate (s%neighbors(iatom)%neigh_j(n))
          allocate (s%neighbors(iatom)%neigh_back(n))
          s%neighbors(iatom)%neigh_j = nj(p+1:p+n)
          s%neighbors(iatom)%neigh_b = nb(p+1:p+n)
          s%neighbors(iatom)%neigh_back = nback(p+1:p+n)
          p = p + n
        end do
        p = 0
        do iatom = 1, s%natoms
          n = nc(iatom)
          s%neighbors(iatom)%ncommon = n
          allocate (s%neighbors(iatom)%iatom_common_b(n))
          allocate (s%neighbors(iatom)%iatom_common_j(n))
          allocate (s%neighbors(iatom)%jatom_common_b(n))
          allocate (s%neighbors(iatom)%jatom_common_j(n))
          allocate (s%neighbors(iatom)%neigh_common(n))
          s%neighbors(iatom)%iatom_common_j = cij(p+1:p+n)
          s%neighbors(iatom)%iatom_common_b = cib(p+1:p+n)
          s%neighbors(iatom)%jatom_common_j = cjj(p+1:p+n)
          s%neighbors(iatom)%jatom_common_b = cjb(p+1:p+n)
          s%neighbors(iatom)%neigh_common = cm(p+1:p+n)
          p = p + n
        end do
        deallocate (nn, nj, nb, nback, nself, nc, cij, cib, cjj, cjb, cm)

        if (iwriteout_neighbors .eq. 1) call writeout_neighbors (s)
        end subroutine driver_neighbors

! The neighbour report of the reference (M_neighbors.f90:839-932): a table
! per atom on s%logfile and the <basis>.NEIGHBORS map, in the same formats.
        subroutine writeout_neighbors (s)
        type(T_structure), target :: s
        integer iatom, ineigh, jatom, mbeta, n
        real(kind=8) :: d(3), r1(3)
        character(len=25) :: fname

        write (s%logfile,*) '  '
        write (s%logfile,*) ' Neighbors of each atom: '
        do iatom = 1, s%natoms
          r1 = s%atom(iatom)%ratom
          n = s%neighbors(iatom)%neighn
          write (s%logfile,*) '  '
          write (s%logfile,'(70(''=''))')
          write (s%logfile,'(2x, '' Atom: '', i4, '','', '' Number of neighbors: '', i4)') iatom, n
          write (s%logfile,'(2x, '' iatom '', '' ineigh '', '' mbeta '', '' jatom '', '' distance '', 10x, '' vector '')')
          write (s%logfile,'(70(''=''))')
          do ineigh = 1, n
            mbeta = s%neighbors(iatom)%neigh_b(ineigh)
            jatom = s%neighbors(iatom)%neigh_j(ineigh)
            d = s%atom(jatom)%ratom + s%xl(mbeta)%a - r1
            write (s%logfile,'(2x, i5, 4x, i3, 5x, i3, 3x, i4, 2x, f9.4, 3f9.4)') iatom, ineigh, mbeta, jatom,  &
     &        sqrt(d(1)**2 + d(2)**2 + d(3)**2), d
          end do
        end do

        fname = s%basisfile(:len(trim(s%basisfile)) - 4)
        fname = trim(fname)//'.NEIGHBORS'
        open (unit = s%inpfile, file = fname, status = 'unknown')
        write (s%inpfile,'(2x, i5, 2x, a40)') s%natoms, s%basisfile
        do iatom = 1, s%natoms
          n = s%neighbors(iatom)%neighn
          write (s%inpfile,*) iatom, n
          do ineigh = 1, n
            write (s%inpfile,*) ineigh, s%neighbors(iatom)%neigh_b(ineigh),                 &
     &        s%neighbors(iatom)%neigh_j(ineigh), s%neigh_self(iatom)
          end do
        end do
        close (unit = s%inpfile)
        end subroutine writeout_neighbors

! Frees what driver_neighbors allocated (M_neighbors.f90:958-1001).
        subroutine destroy_neighbors (s)
        type(T_structure), target :: s
        integer iatom

        do iatom = 1, s%natoms
          deallocate (s%neighbors(iatom)%iatom_common_b, s%neighbors(iatom)%iatom_common_j)
          deallocate (s%neighbors(iatom)%jatom_common_b, s%neighbors(iatom)%jatom_common_j)
          deallocate (s%neighbors(iatom)%neigh_common)
          deallocate (s%neighbors(iatom)%neigh_j, s%neighbors(iatom)%neigh_b)
          deallocate (s%neighbors(iatom)%neigh_back)
        end do
        deallocate (s%neighbors)
        deallocate (s%neigh_self)
        end subroutine destroy_neighbors

        end module M_neighbors
