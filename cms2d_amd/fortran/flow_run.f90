!=============================================================================
! flow_run.f90 -- Fortran host program: the CMS_Flow time loop
! (source/main/CMS_Flow.f90:43-87) over the device-resident implicit step.
!
! Usage: flow_run <case file> <nsteps> <output file>
!        flow_run --sizeof   (prints the sizes of the two bind(C) structs)
!   case file (float64 stream): nx, ny, dx, dy, dt, amp, speed[rad/hr],
!   phase[rad], hdry, noptset, then zb(nx*ny), mann(nx*ny) and, if noptset>=3,
!   wavestrx(nx*ny), wavestry(nx*ny).  East side harmonic water level, walls
!   elsewhere.  Writes eta, u, v (float64 stream, nx*ny each) after nsteps and
!   prints one line per step (outer iterations, Krylov iterations).
!=============================================================================
program flow_run
    use iso_c_binding
    use cms2d_hip
    implicit none
    character(len=400) :: cfile, arg, ofile
    integer :: nsteps, n, u, nc, rc
    real(8) :: hdr(10)
    real(8), allocatable, target :: zb(:), mann(:), wx(:), wy(:)
    real(8), allocatable :: eta(:), uu(:), vv(:)
    type(cms2d_flow_desc) :: d
    type(cms2d_step_info) :: info
    type(c_ptr) :: flow

    call get_command_argument(1, cfile)
    if (trim(cfile) == '--sizeof') then
        write(*, '(i0,1x,i0)') c_sizeof(d), c_sizeof(info)
        stop
    end if
    call get_command_argument(2, arg)
    read(arg, *) nsteps
    call get_command_argument(3, ofile)

    open(newunit=u, file=trim(cfile), access='stream', form='unformatted', status='old')
    read(u) hdr
    d%nx = nint(hdr(1)); d%ny = nint(hdr(2)); d%dx = hdr(3); d%dy = hdr(4); d%dt = hdr(5)
    d%ntc = 1; d%amp(1) = hdr(6); d%speed(1) = hdr(7); d%phase(1) = hdr(8)
    d%hdry = hdr(9); d%noptset = nint(hdr(10))
    nc = d%nx * d%ny
    allocate(zb(nc), mann(nc), eta(nc), uu(nc), vv(nc))
    read(u) zb
    read(u) mann
    d%zb = c_loc(zb)
    d%mann = c_loc(mann)
    if (d%noptset >= 3) then
        allocate(wx(nc), wy(nc))
        read(u) wx
        read(u) wy
        d%wavestrx = c_loc(wx)
        d%wavestry = c_loc(wy)
    end if
    close(u)
    d%side_type = [CMS2D_BND_WALL, CMS2D_BND_WSE_HARMONIC, CMS2D_BND_WALL, CMS2D_BND_WALL]

    rc = cms2d_flow_create(d, flow)
    if (rc /= CMS2D_OK) then
        write(*, '(a,i0,2a)') 'cms2d_flow_create failed: ', rc, ' ', trim(cms2d_error_message())
        stop 2
    end if
    do n = 1, nsteps
        rc = cms2d_flow_step(flow, 1, info)
        if (rc /= CMS2D_OK) then
            write(*, '(a,i0,2a)') 'cms2d_flow_step failed: ', rc, ' ', trim(cms2d_error_message())
            stop 3
        end if
        write(*, '(a,i0,a,i0,a,3i7)') 'step ', n, ' niter ', info%niter, ' krylov ', info%kry_iters
    end do
    rc = cms2d_flow_get(flow, 'eta', eta)
    rc = cms2d_flow_get(flow, 'u', uu)
    rc = cms2d_flow_get(flow, 'v', vv)
    rc = cms2d_flow_destroy(flow)
    open(newunit=u, file=trim(ofile), access='stream', form='unformatted', status='replace')
    write(u) eta
    write(u) uu
    write(u) vv
    close(u)
end program flow_run
