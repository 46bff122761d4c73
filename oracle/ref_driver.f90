!=============================================================================
! ref_driver.f90 -- runs the reference's own implicit flow step on an ASCII
! case and writes the state after every step.
!
! TEST INFRASTRUCTURE ONLY: a checker for the FP64 oracle.  Built by
! oracle/build_ref.py against the reference's objects (compiled unmodified,
! single precision as shipped) into oracle/_ref/ref_run.
!
! Usage: ref_run <case dir/> <case name> <nsteps> <output file|-> [forcing file]
!   reads <case dir>/<case name>.cmcards (and the files it names: a .cart or
!   a .tel grid) with the reference's card readers, initialises the model with the calls prestart
!   makes (source/main/prestart.f90:2-329) except those of the output module
!   (source/output/output.F90 needs Intel's IFPORT module and is not built;
!   it writes result files only), then runs the time loop of CMS_Flow
!   (source/main/CMS_Flow.f90:43-87): timing, flow_imp, flow_update.  After
!   each step appends, as float64 stream data: ncells, then eta, u, v, iwet.
!   The optional forcing file (float64 stream, 2*ncells values) replaces the
!   radiation-stress arrays wavestrx, wavestry after initialisation; it is
!   used with the CONSTANT_WAVE_HEIGHT card, which turns the wave branches on
!   with zero wave height (source/waves/wave.F90:506-540).  With output file
!   "-" nothing is written.  Prints the wall time of every step (timing,
!   flow_imp, flow_update; output excluded) as "step_seconds <n> <t>".
!   With the environment variable CMS2D_REF_BALANCE naming a file (and the
!   card WATER_VOLUME_BALANCE ON, so that flow_imp calls flow_balance,
!   source/flow/flow_imp.F90:305) every step also appends, as float64 stream
!   data, the cumulative volumes of flow_balance (volH2Ocumstg, volH2Ocumbnd,
!   volH2Ocumflux, volH2Ocumnorm; flow_imp.F90:2260-2372) and the extremes of
!   flow_step_stat (pxtrm, idp, uxtrm, idu, vxtrm, idv, crmax, idcr, cell
!   numbers 1-based; source/flow/flow.F90:1912-2008).
!   With CMS2D_REF_STATS naming a file (and a FLOW_STATISTICS card) the loop
!   also calls stat_update where CMS_Flow does (after flow_imp, before
!   flow_update, source/main/CMS_Flow.f90:78-81) and appends after each step
!   nflowstatcount and the flow statistics of every cell (stat_def, in the
!   order of STATS below).
!=============================================================================
program ref_run
    use comvarbl,  only: ctlfile, flowpath, casename
    use cms_def,   only: cmsflow, cmswave, noptset
    use diag_def,  only: dgfile, dgunit, interactive
    use geo_def,   only: telfile
    use size_def,  only: ncells
    use flow_def,  only: eta, u, v, iwet, volH2Ocumstg, volH2Ocumbnd, volH2Ocumflux, volH2Ocumnorm
    use comvarbl,  only: pxtrm, idp, uxtrm, idu, vxtrm, idv, crmax, idcr
    use stat_def,  only: nflowstatcount, etamax, tetamax, etamin, tetamin, uvmax, umax, vmax, tuvmax, &
        etamean, ur, vr, uvcurv, hydper, rsuavg, rsvavg, rspavg, rsumax, rsvmax, rspmax, etamaxgrad, &
        uvmaxgrad, frxmax, frymax, umaxcurvx, umaxcurvy, vmaxcurvx, vmaxcurvy, cflxmax, cflymax, &
        qcflxmax, qcflymax, qFrxmax, qFrymax
    use wave_flowgrid_def, only: wavestrx, wavestry, constant_waves
    implicit none
    character(len=400) :: dir, name, arg, outfile, forcefile, balfile, statfile
    integer :: nsteps, n, i, ou, fu, bu, blen, su, slen
    integer(8) :: t0, t1, rate
    logical :: dump
    real(8), allocatable :: buf(:)
    interface
      subroutine model_init()
      end subroutine
    end interface

    call get_command_argument(1, dir)
    call get_command_argument(2, name)
    call get_command_argument(3, arg)
    read(arg, *) nsteps
    call get_command_argument(4, outfile)

    ! the setup get_com_arg does (source/main/CMS2D_main.f90:110-390)
    interactive = .false.
    cmsflow = .true.
    cmswave = .false.
    noptset = 2
    flowpath = trim(dir)
    casename = trim(name)
    ctlfile = trim(dir) // trim(name) // '.cmcards'
    telfile = ' '
    dgfile = trim(dir) // 'CMS_DIAG.txt'
    dgunit = 9
    open(dgunit, file=dgfile, status='unknown')
    close(dgunit)

    call model_init

    forcefile = ' '
    if (command_argument_count() >= 5) call get_command_argument(5, forcefile)
    if (len_trim(forcefile) > 0) then
      allocate(buf(2*ncells))
      open(newunit=fu, file=trim(forcefile), access='stream', form='unformatted', status='old')
      read(fu) buf
      close(fu)
      do i = 1, ncells
        wavestrx(i) = buf(i)
        wavestry(i) = buf(ncells + i)
      enddo
      deallocate(buf)
    endif

    call get_environment_variable('CMS2D_REF_STATS', statfile, slen)
    if (slen > 0) open(newunit=su, file=trim(statfile), access='stream', form='unformatted', status='replace')
    call get_environment_variable('CMS2D_REF_BALANCE', balfile, blen)
    if (blen > 0) open(newunit=bu, file=trim(balfile), access='stream', form='unformatted', status='replace')
    dump = trim(outfile) /= '-'
    if (dump) open(newunit=ou, file=trim(outfile), access='stream', form='unformatted', status='replace')
    allocate(buf(ncells))
    do n = 1, nsteps
      call system_clock(t0, rate)
      call timing
      if (constant_waves) call wave_const_update
      call flow_imp
      if (slen > 0) call stat_update
      if (slen > 0) then
        write(su) real(nflowstatcount, 8)
        write(su) (real(etamax(i), 8), i = 1, ncells), (real(tetamax(i), 8), i = 1, ncells)
        write(su) (real(etamin(i), 8), i = 1, ncells), (real(tetamin(i), 8), i = 1, ncells)
        write(su) (real(uvmax(i), 8), i = 1, ncells), (real(umax(i), 8), i = 1, ncells)
        write(su) (real(vmax(i), 8), i = 1, ncells), (real(tuvmax(i), 8), i = 1, ncells)
        write(su) (real(etamean(i), 8), i = 1, ncells), (real(ur(i), 8), i = 1, ncells)
        write(su) (real(vr(i), 8), i = 1, ncells), (real(uvcurv(i), 8), i = 1, ncells)
        write(su) (real(hydper(i), 8), i = 1, ncells), (real(rsuavg(i), 8), i = 1, ncells)
        write(su) (real(rsvavg(i), 8), i = 1, ncells), (real(rspavg(i), 8), i = 1, ncells)
        write(su) (real(rsumax(i), 8), i = 1, ncells), (real(rsvmax(i), 8), i = 1, ncells)
        write(su) (real(rspmax(i), 8), i = 1, ncells), (real(etamaxgrad(i), 8), i = 1, ncells)
        write(su) (real(uvmaxgrad(i), 8), i = 1, ncells), (real(frxmax(i), 8), i = 1, ncells)
        write(su) (real(frymax(i), 8), i = 1, ncells), (real(umaxcurvx(i), 8), i = 1, ncells)
        write(su) (real(umaxcurvy(i), 8), i = 1, ncells), (real(vmaxcurvx(i), 8), i = 1, ncells)
        write(su) (real(vmaxcurvy(i), 8), i = 1, ncells), (real(cflxmax(i), 8), i = 1, ncells)
        write(su) (real(cflymax(i), 8), i = 1, ncells), (real(qcflxmax(i), 8), i = 1, ncells)
        write(su) (real(qcflymax(i), 8), i = 1, ncells), (real(qFrxmax(i), 8), i = 1, ncells)
        write(su) (real(qFrymax(i), 8), i = 1, ncells)
      endif
      call flow_update
      call system_clock(t1)
      write(*, '(a,i0,1x,es14.6)') 'step_seconds ', n, real(t1 - t0, 8) / real(rate, 8)
      if (blen > 0) write(bu) real(volH2Ocumstg, 8), real(volH2Ocumbnd, 8), real(volH2Ocumflux, 8), &
          real(volH2Ocumnorm, 8), real(pxtrm, 8), real(idp, 8), real(uxtrm, 8), real(idu, 8), &
          real(vxtrm, 8), real(idv, 8), real(crmax, 8), real(idcr, 8)
      if (.not. dump) cycle
      write(ou) real(ncells, 8)
      do i = 1, ncells
        buf(i) = real(eta(i), 8)
      enddo
      write(ou) buf
      do i = 1, ncells
        buf(i) = real(u(i), 8)
      enddo
      write(ou) buf
      do i = 1, ncells
        buf(i) = real(v(i), 8)
      enddo
      write(ou) buf
      do i = 1, ncells
        buf(i) = real(iwet(i), 8)
      enddo
      write(ou) buf
    enddo
    if (dump) close(ou)
    if (blen > 0) close(bu)
    if (slen > 0) close(su)
end program ref_run

!-----------------------------------------------------------------------------
! The initialisation prestart performs (source/main/prestart.f90:63-329),
! without the output module's routines, the hot start and the printouts.
!-----------------------------------------------------------------------------
subroutine model_init()
    use comvarbl
    use cms_def
    use geo_def
    use geo_lib,    only: assign_proj_names
    use flow_def
    use der_def
    use der_lib,    only: der_grad_eval
    use interp_lib, only: interp_scal_cell2face
    use size_def
#ifdef REF_NO_DREDGE
    use dredge_def, only: dredging
#endif
    implicit none
    integer :: ierr
    character(len=37) :: cardname
    logical :: foundcard

    call geo_default
    call flow_default
    call solv_default
    call fric_default
    call bnd_default
    call struct_default
    call hot_default
    call sed_default
    call sal_default
    call heat_default
    call met_default
    call rol_default
    call stat_default
    call der_default
    call diag_default
    call wave_default
#ifdef REF_NO_DREDGE
    ! the double-precision build leaves out the dredging module's setup
    ! (its HDF5 reads do not compile at ikind=8); dredging stays off
    dredging = .false.
#else
    call dredge_default
#endif

    open(77, file=ctlfile, status='old')
    do
      read(77, *, iostat=ierr) cardname
      if (ierr /= 0) exit
      if (cardname(1:14) == 'END_PARAMETERS') exit
      if (cardname(1:1) == '!' .or. cardname(1:1) == '#' .or. cardname(1:1) == '*') cycle
      call ignr_cards(cardname, foundcard);          if (foundcard) cycle
      call geo_cards(cardname, foundcard, .true.);   if (foundcard) cycle
      call flow_cards(cardname, foundcard, .true.);  if (foundcard) cycle
      call bnd_cards(cardname, foundcard);           if (foundcard) cycle
      call struct_cards(cardname, foundcard);        if (foundcard) cycle
      call fric_cards(cardname, foundcard);          if (foundcard) cycle
      call hot_cards(cardname, foundcard);           if (foundcard) cycle
      call sed_cards(cardname, foundcard);           if (foundcard) cycle
      call sal_cards(cardname, foundcard);           if (foundcard) cycle
      call heat_cards(cardname, foundcard);          if (foundcard) cycle
      call met_cards(cardname, foundcard);           if (foundcard) cycle
      call rol_cards(cardname, foundcard);           if (foundcard) cycle
      call diag_cards(cardname, foundcard);          if (foundcard) cycle
      call wave_cards(cardname, foundcard);          if (foundcard) cycle
#ifndef REF_NO_DREDGE
      call dredge_cards(cardname, foundcard);        if (foundcard) cycle
#endif
      call der_cards(cardname, foundcard);           if (foundcard) cycle
      call stat_cards(cardname, foundcard)
    enddo
    close(77)
    call assign_proj_names(projfl)

    ! the grid file prestart reads (prestart.f90:190-218): a telescoping grid
    ! (.tel, TELESCOPING card) or a Cartesian one (.cart)
    if (len_trim(telfile) > 0) then
      igridtype = 1
      call read_tel
    else
      call read_grid_cart
    endif
    call geo_init
    call bnd_init
    call struct_init
    call solv_init
    call flow_init
    call interp_init
    call der_init
    call fric_init
    call met_init
    call stat_init
    call wave_init

    call flow_wetdry(0)
    if (sum(1 - iwet(1:ncells)) == ncells) then
      ponding = .true.
      call flow_wetdry(0)
    endif
    iwet1 = iwet
    eta = iwet*p*gravinv - 999.0*(1 - iwet)
    call der_grad_eval(goa, 0, zb, dzbx, dzby)
    call interp_scal_cell2face(zb, 0, zbk, dzbx, dzby)
    call flow_grad_interp
    call flow_convflux
    call fric_eval
    call flow_eddyvis
end subroutine model_init
