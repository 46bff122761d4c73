!=============================================================================
! solve_run.f90 -- Fortran caller of cms2d_solve_ref with the reference's
! array shapes: the call solve(acoef,ss,sp,res,phi,n) makes in flow_imp,
! sed_imp, salinity and heat (source/solvers/solvers.f90:205-304), with
! acoef(nmaxfaces,ncellsD) and 1-based cell2cell(nmaxfaces,ncellsD) as geo_def
! holds them (source/geospatial/geo_def.f90:22-157).
!
! Usage: solve_run <system file> <output file>
!   system file (stream): int32 ncells, ncellsD, nmaxfaces; float64 relax;
!   int32 cell2cell(nmaxfaces,ncellsD) (1-based, 0 = no face), iwet(ncellsD);
!   float64 acoef(nmaxfaces,ncellsD), ss, sp, phi (ncellsD each).
!   Writes float64 phi(ncellsD), res(ncellsD), ss(ncellsD), rmom and prints
!   the status and the Krylov iterations.
!=============================================================================
program solve_run
    use iso_c_binding
    use cms2d_hip
    implicit none
    character(len=400) :: sfile, ofile
    integer(4) :: ncells, ncellsD, nmaxfaces
    integer :: u, rc, iters
    real(8) :: relax, rmom
    integer(4), allocatable :: cell2cell(:,:), iwet(:)
    real(8), allocatable :: acoef(:,:), ss(:), sp(:), phi(:), res(:)

    call get_command_argument(1, sfile)
    call get_command_argument(2, ofile)
    open(newunit=u, file=trim(sfile), access='stream', form='unformatted', status='old')
    read(u) ncells, ncellsD, nmaxfaces
    read(u) relax
    allocate(cell2cell(nmaxfaces, ncellsD), iwet(ncellsD), acoef(nmaxfaces, ncellsD))
    allocate(ss(ncellsD), sp(ncellsD), phi(ncellsD), res(ncellsD))
    read(u) cell2cell, iwet, acoef, ss, sp, phi
    close(u)

    rc = cms2d_solve_ref(ncells, ncellsD, nmaxfaces, acoef, cell2cell, iwet, ss, sp, res, phi, relax, &
                         rmom, iters)
    write(*, '(a,i0,a,i0)') 'status ', rc, ' iterations ', iters
    if (rc /= 0) then
        write(*, '(a)') trim(cms2d_error_message())
        stop 1
    end if
    open(newunit=u, file=trim(ofile), access='stream', form='unformatted', status='replace')
    write(u) phi, res, ss, rmom
    close(u)
end program solve_run
