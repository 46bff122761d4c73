!=============================================================================
! cms2d_hip_mod.f90 -- Fortran host layer of the MI355X implicit CMS-Flow step.
!
! ISO_C_BINDING interfaces to the C ABI in include/cms2d_amd.h and the
! reference-shaped wrappers a Fortran caller uses:
!   cms2d_solve_ref   -- solve(acoef,ss,sp,res,phi,n) on the reference's cell
!                        list (source/solvers/solvers.f90:205-304), with the
!                        reference's array shapes acoef(nmaxfaces,ncellsD),
!                        1-based cell2cell(nmaxfaces,ncellsD) and iwet(ncellsD)
!   cms2d_flow_*      -- the whole implicit step (flow_imp,
!                        source/flow/flow_imp.F90:29-312) kept on the device,
!                        driven from a Fortran time loop as CMS_Flow does
!                        (source/main/CMS_Flow.f90:43-87).
!=============================================================================
module cms2d_hip
    use iso_c_binding
    implicit none
    private

    integer(c_int), parameter, public :: CMS2D_OK = 0, CMS2D_ERR_ARG = -1, CMS2D_ERR_HIP = -2, &
                                         CMS2D_ERR_DIVERGED = -3, CMS2D_ERR_NODEVICE = -4
    integer(c_int), parameter, public :: CMS2D_BND_WALL = 0, CMS2D_BND_WSE_HARMONIC = 1, CMS2D_BND_WSE_CURVE = 2, &
                                         CMS2D_BND_FLUX = 3
    integer(c_int), parameter, public :: CMS2D_PREC_FP64 = 0
    integer(c_int), parameter, public :: CMS2D_PRECOND_MG = 0, CMS2D_PRECOND_JACOBI = 1
    integer(c_int), parameter, public :: CMS2D_SOLVER_BICGSTAB = 0, CMS2D_SOLVER_GMRES = 1
    integer(c_int), parameter, public :: CMS2D_TURB_SUBGRID = 0, CMS2D_TURB_CONSTANT = 1, CMS2D_TURB_FALCONER = 2, &
                                         CMS2D_TURB_PARABOLIC = 3, CMS2D_TURB_SUBGRID_WU = 4
    integer(c_int), parameter, public :: CMS2D_FRIC_MANNING = 0, CMS2D_FRIC_DRAG = 1, CMS2D_FRIC_ROUGHNESS = 2, &
                                         CMS2D_FRIC_LINEAR = 3

    ! struct cms2d_flow_desc
    type, bind(C), public :: cms2d_flow_desc
        integer(c_int) :: nx = 0, ny = 0
        real(c_double) :: dx = 0, dy = 0
        type(c_ptr) :: zb = c_null_ptr, mann = c_null_ptr, wavestrx = c_null_ptr, wavestry = c_null_ptr
        integer(c_int) :: side_type(4) = 0
        integer(c_int) :: ntc = 0
        real(c_double) :: amp(8) = 0, speed(8) = 0, phase(8) = 0
        real(c_double) :: dt = 0, rampdur = 0, hdry = 0
        integer(c_int) :: noptset = 2, wallfric = 1, ndsch = 4, maxit = 0
        real(c_double) :: rtol = 1.0d-10, atol = 1.0d-12
        integer(c_int) :: kmax = 4000, precision = 0, device = 0
        integer(c_int) :: precond = 0, solver = 0, dtvar = 0
        real(c_double) :: latitude = 0
        integer(c_int) :: coriolis_plane = 0
        integer(c_int) :: turbulence = 0
        integer(c_int) :: friction = 0
        type(c_ptr) :: wse_times = c_null_ptr, wse_values = c_null_ptr
        integer(c_int) :: wse_ntimes = 0, wse_order = 0
        integer(c_int) :: wse_adjust_off = 0
        real(c_double) :: wse_gradx = 0, wse_grady = 0
        integer(c_int) :: serial_order = 0
        real(c_double) :: flux_q = 0                                   ! CMS2D_BND_FLUX side: inflow [m^3/s]
        type(c_ptr) :: flux_times = c_null_ptr, flux_values = c_null_ptr   ! or its curve [hr], [m^3/s]
        integer(c_int) :: flux_ntimes = 0, flux_order = 0
        real(c_double) :: flux_conveyance = 0
        integer(c_int) :: flux_angle_on = 0                            ! INFLOW_ANGLE given
        real(c_double) :: flux_angle = 0                               ! [deg clockwise from north]
        integer(c_int) :: flux_rating = 0                              ! RATING_CURVE: stage [m] -> flow
    end type

    ! struct cms2d_step_info
    type, bind(C), public :: cms2d_step_info
        integer(c_int) :: niter = 0, iconv = 0
        real(c_double) :: rmom(3) = 0
        integer(c_int) :: kry_iters(3) = 0
        integer(c_int) :: nwet = 0
        real(c_double) :: kernel_ms = 0, dt = 0
    end type

    ! struct cms2d_balance: flow_balance and flow_step_stat of the last step
    type, bind(C), public :: cms2d_balance
        real(c_double) :: storage = 0, efflux = 0, flux = 0, balance = 0, net = 0, error_pct = 0
        real(c_double) :: cum_storage = 0, cum_efflux = 0, cum_flux = 0, cum_balance = 0, cum_net = 0, &
                          cum_error_pct = 0
        real(c_double) :: wse = 0, u = 0, v = 0, courant = 0
        integer(c_int) :: wse_cell = 0, u_cell = 0, v_cell = 0, courant_cell = 0   ! 0-based
        integer(c_int) :: steps = 0
    end type

    ! struct cms2d_tel_grid: the content of a .tel file (read_tel)
    type, bind(C), public :: cms2d_tel_grid
        integer(c_int) :: ncells = 0
        type(c_ptr) :: geom = c_null_ptr        ! real(c_double) geom(4, ncells): x, y, dx, dy
        type(c_ptr) :: nb = c_null_ptr          ! integer(c_int) nb(8, ncells): N1 N2 E1 E2 S1 S2 W1 W2
        type(c_ptr) :: bnd_cells = c_null_ptr   ! integer(c_int) bnd_cells(nbnd), 0-based
        integer(c_int) :: nbnd = 0
        type(c_ptr) :: flux_cells = c_null_ptr  ! integer(c_int) flux_cells(nflux), 0-based
        integer(c_int) :: nflux = 0
    end type

    interface
        function cms2d_flow_create_tel(desc, grid, flow) bind(C, name='cms2d_flow_create_tel') result(rc)
            import :: c_int, c_ptr, cms2d_flow_desc, cms2d_tel_grid
            type(cms2d_flow_desc), intent(in) :: desc
            type(cms2d_tel_grid), intent(in) :: grid
            type(c_ptr), intent(out) :: flow
            integer(c_int) :: rc
        end function
        function cms2d_last_error() bind(C, name='cms2d_last_error') result(p)
            import :: c_ptr
            type(c_ptr) :: p
        end function
        function cms2d_device_count(n) bind(C, name='cms2d_device_count') result(rc)
            import :: c_int
            integer(c_int), intent(out) :: n
            integer(c_int) :: rc
        end function
        function c_flow_create(desc, flow) bind(C, name='cms2d_flow_create') result(rc)
            import :: c_int, c_ptr, cms2d_flow_desc
            type(cms2d_flow_desc), intent(in) :: desc
            type(c_ptr), intent(out) :: flow
            integer(c_int) :: rc
        end function
        function c_flow_destroy(flow) bind(C, name='cms2d_flow_destroy') result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: flow
            integer(c_int) :: rc
        end function
        function c_flow_step(flow, nsteps, info) bind(C, name='cms2d_flow_step') result(rc)
            import :: c_int, c_ptr, cms2d_step_info
            type(c_ptr), value :: flow
            integer(c_int), value :: nsteps
            type(cms2d_step_info), intent(out) :: info
            integer(c_int) :: rc
        end function
        function c_flow_get(flow, name, out) bind(C, name='cms2d_flow_get') result(rc)
            import :: c_int, c_ptr, c_char, c_double
            type(c_ptr), value :: flow
            character(kind=c_char), intent(in) :: name(*)
            real(c_double), intent(out) :: out(*)
            integer(c_int) :: rc
        end function
        function c_solve(ncells, ncellsD, nmaxfaces, acoef, c2c, iwet, ss, sp, res, phi, relax, rtol, atol, &
                         kmax, rmom, iters) bind(C, name='cms2d_solve') result(rc)
            import :: c_int, c_double
            integer(c_int), value :: ncells, ncellsD, nmaxfaces, kmax
            real(c_double), intent(in) :: acoef(*), sp(*)
            integer(c_int), intent(in) :: c2c(*), iwet(*)
            real(c_double), intent(inout) :: ss(*), phi(*)
            real(c_double), intent(out) :: res(*)
            real(c_double), value :: relax, rtol, atol
            real(c_double), intent(out) :: rmom
            integer(c_int), intent(out) :: iters
            integer(c_int) :: rc
        end function
        ! hot start, exact-resume variant (hot_start.F90:1512-1560 write, :693-1344 read)
        function cms2d_flow_checkpoint(flow, buf, nbytes) bind(C, name='cms2d_flow_checkpoint') result(rc)
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: flow, buf
            integer(c_size_t), intent(inout) :: nbytes
            integer(c_int) :: rc
        end function
        function cms2d_flow_restore(flow, buf, nbytes) bind(C, name='cms2d_flow_restore') result(rc)
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: flow, buf
            integer(c_size_t), value :: nbytes
            integer(c_int) :: rc
        end function
        ! windcurve_eval, spatially constant wind (meteorological.F90:995-1033)
        function cms2d_flow_set_wind(flow, wndx, wndy) bind(C, name='cms2d_flow_set_wind') result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: flow
            real(c_double), value :: wndx, wndy
            integer(c_int) :: rc
        end function
        ! wave_rad_stress_fl (waves/wave_rad_stress_fl.f90:2-59)
        function cms2d_flow_set_waves(flow, whgt, wper, wlen, wunitx, wunity) &
                bind(C, name='cms2d_flow_set_waves') result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: flow
            real(c_double), intent(in) :: whgt(*), wper(*), wlen(*), wunitx(*), wunity(*)
            integer(c_int) :: rc
        end function
        ! flow_balance (flow_imp.F90:2260-2372) and flow_step_stat (flow.F90:1912-2008)
        function cms2d_flow_balance(flow, bal) bind(C, name='cms2d_flow_balance') result(rc)
            import :: c_int, c_ptr, cms2d_balance
            type(c_ptr), value :: flow
            type(cms2d_balance), intent(out) :: bal
            integer(c_int) :: rc
        end function
        ! discrete adjoint of the time loop: dL/dn and dL/dzb of a gauge water-level loss
        function cms2d_flow_adjoint(flow, nsteps, every, ngauge, gauge_cells, gauge_obs, geta, eta_gauge, &
                                    loss, gmann, gzb) bind(C, name='cms2d_flow_adjoint') result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: flow
            integer(c_int), value :: nsteps, every, ngauge
            integer(c_int), intent(in) :: gauge_cells(*)
            type(c_ptr), value :: gauge_obs, geta, eta_gauge, loss
            real(c_double), intent(out) :: gmann(*), gzb(*)
            integer(c_int) :: rc
        end function
    end interface

    public :: cms2d_solve_ref, cms2d_flow_create, cms2d_flow_step, cms2d_flow_get, cms2d_flow_destroy
    public :: cms2d_error_message, cms2d_device_count
    public :: cms2d_flow_checkpoint, cms2d_flow_restore, cms2d_flow_set_waves, cms2d_flow_set_wind
    public :: cms2d_flow_adjoint, cms2d_flow_balance
    public :: cms2d_flow_create_tel

contains

    function cms2d_error_message() result(msg)
        character(len=256) :: msg
        character(kind=c_char), pointer :: s(:)
        type(c_ptr) :: p
        integer :: i
        msg = ' '
        p = cms2d_last_error()
        if (.not. c_associated(p)) return
        call c_f_pointer(p, s, [256])
        do i = 1, 256
            if (s(i) == c_null_char) exit
            msg(i:i) = s(i)
        end do
    end function

    ! solve(acoef,ss,sp,res,phi,n) with the reference's shapes.  cell2cell is
    ! 1-based as in geo_def (0 = no face); relax = relax(n) of comvarbl.
    ! Returns the status; rmom = sqrt(adak(res)/ncells), iters = Krylov iterations.
    function cms2d_solve_ref(ncells, ncellsD, nmaxfaces, acoef, cell2cell, iwet, ss, sp, res, phi, relax, &
                             rmom, iters, rtol) result(rc)
        integer, intent(in) :: ncells, ncellsD, nmaxfaces
        real(8), intent(in) :: acoef(nmaxfaces, ncellsD), sp(ncellsD)
        integer, intent(in) :: cell2cell(nmaxfaces, ncellsD), iwet(ncellsD)
        real(8), intent(inout) :: ss(ncellsD), phi(ncellsD)
        real(8), intent(out) :: res(ncellsD)
        real(8), intent(in) :: relax
        real(8), intent(out) :: rmom
        integer, intent(out) :: iters
        real(8), intent(in), optional :: rtol
        integer(c_int) :: rc
        integer(c_int), allocatable :: c2c(:)
        real(c_double), allocatable :: a4(:)
        integer(c_int), allocatable :: iw(:)
        integer :: i, k
        real(c_double) :: tol
        integer(c_int) :: it
        allocate(c2c(nmaxfaces*ncellsD), a4(nmaxfaces*ncellsD), iw(ncellsD))
        do i = 1, ncellsD
            do k = 1, nmaxfaces
                c2c(nmaxfaces*(i-1)+k) = cell2cell(k, i) - 1         ! 0-based, -1 for none
                a4(nmaxfaces*(i-1)+k) = acoef(k, i)
            end do
            iw(i) = iwet(i)
        end do
        tol = 1.0d-10
        if (present(rtol)) tol = rtol
        rc = c_solve(int(ncells, c_int), int(ncellsD, c_int), int(nmaxfaces, c_int), a4, c2c, iw, ss, sp, res, phi, &
                     real(relax, c_double), tol, 1.0d-12, 4000_c_int, rmom, it)
        iters = it
    end function

    function cms2d_flow_create(desc, flow) result(rc)
        type(cms2d_flow_desc), intent(in) :: desc
        type(c_ptr), intent(out) :: flow
        integer(c_int) :: rc
        rc = c_flow_create(desc, flow)
    end function

    function cms2d_flow_step(flow, nsteps, info) result(rc)
        type(c_ptr), intent(in) :: flow
        integer, intent(in) :: nsteps
        type(cms2d_step_info), intent(out) :: info
        integer(c_int) :: rc
        rc = c_flow_step(flow, int(nsteps, c_int), info)
    end function

    function cms2d_flow_get(flow, name, out) result(rc)
        type(c_ptr), intent(in) :: flow
        character(len=*), intent(in) :: name
        real(8), intent(out) :: out(*)
        integer(c_int) :: rc
        rc = c_flow_get(flow, trim(name)//c_null_char, out)
    end function

    function cms2d_flow_destroy(flow) result(rc)
        type(c_ptr), intent(inout) :: flow
        integer(c_int) :: rc
        rc = c_flow_destroy(flow)
        flow = c_null_ptr
    end function

end module cms2d_hip
