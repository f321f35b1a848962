! mgcm_rt_driver.f90 -- a COMP3-style host program over mgcm_rt_mod: reads a
! radiation case (tables and a batch of columns) from a stream file, sets the
! tables up once, runs the batch through mgcm_rt_radiation and writes the
! results to a second stream file.
!
!   mgcm_rt_driver [--per-column] [--aerosol <tracer file> <aerosol data dir>]
!                  [--dust <dust file>] [--tile <band slots>]
!                  <case file> <result file> [data directory | -] [history file]
!
! --tile sets the column kernel's tile (mgcm_rt_set_tile: 1, 2 or 4 band
! slots per column, 12 = one band per wave, 0 = the library's choice, the
! default).
!
! --dust replaces the case file's TAUREF by the prescribed (Conrath) dust of
! DUSTPROFILE + FILLTAUCUM (mgcm_rt_dust_profile).  Dust file (FP64 stream):
! PSF, PTROP, RPTAU, then TAUTOT(ncol), CONR(ncol), P(ncol), SIGMA(L_LEVELS).
! --aerosol runs COMP3's active dust and water-ice clouds (code/comp3.f:595-631):
! INITCLD's tables from the four optical-property files in <aerosol data dir>,
! then OPT_DST and OPT_CLD on the tracers QTRACE(nlay, 6, ncol) (FP64 stream
! file) and the columns' PLEV; their TAUREF/TAUREFCLD replace the case file's
! and their per-level optics go to the radiation through its aerosol argument.
!
! --per-column runs COMP3's own loop shape instead of one batch: mgcm_rt_column
! once per column, with DETAU/CUMTAUV as the module arrays (L_J, L_I, 7, 17),
! here (ncol, 1, 7, 17), in FP64; the heating is not formed (written as zeros).
! The fluxes, DETAU and CUMTAUV are the batch's, bit for bit.
!
! With a data directory, the k-tables are set up as RADSETUP does on the GCM's
! start (code/setrad.f90:112-122, code/laginterp.f90): CO2H2O_V_12_95_INTEL and
! CO2H2O_IR_12_95_INTEL are read and refined to the 51 PINT points, and CO2V,
! CO2I, PFGASREF, FZEROV and FZEROI of the case file are replaced.
!
! Case file (native-endian stream): int32 ncol, nlay, precision; then FP64
! CO2V(7,51,10,7,17), CO2I(7,51,10,5,17), PFGASREF(51), TGASREF(7),
! WREFH2O(10), GWV(17,7), GWI(17,5), FZEROV(7), FZEROI(5), TAURAY(7), DWNI(5),
! PLANCKIR(5,8501), dust Qext/Qscat/g VIS(7) and IR(5), cloud Qext/Qscat/g
! VIS(7) and IR(5), QEXTREFDST, QEXTREFCLD, SOL(7); then per column set
! PLEV, TLEV, PMID, TMID, QH2O (L_LEVELS, ncol), TAUREF, TAUREFCLD
! (L_LEVELS+1, ncol), ALS, ALBI, ACOSZ (ncol).
! With a history file, the radiation's two history records are written to it
! the way COMP3 fills them (REAL*4 assignments, code/comp3.f:703-706,745-747)
! and MHISTV writes them (Fortran sequential unformatted, code/mhistv.f:200-201),
! the columns taken as (L_J, L_I) = (ncol, 1).
! Result file: int32 status; FMNETV, FLUXUPV, FLUXDNV, FMNETI, FLUXUPI,
! FLUXDNI (L_NLAYRAD, ncol), NFLUXTOPV, NFLUXTOPI, DIFFVT (ncol), DETAU,
! CUMTAUV (7, 17, ncol), heating (L_NLAYRAD, ncol).
program mgcm_rt_driver
  use iso_c_binding
  use mgcm_rt_mod
  implicit none
  character(len=4096) :: fin, fout, datadir, fhist
  ! PGASREF: REAL(4) literals, as SETRAD writes them (code/setrad.f90:112-122)
  real(c_double), parameter :: pgref(11) = [1.0e-6, 1.0e-5, 1.0e-4, 1.0e-3, 1.0e-2, 1.0e-1, 1.0, 1.0e1, &
                                            1.0e2, 1.0e3, 1.0e4]
  integer :: u, ncol, nlay, nlev, nl, ios
  integer(c_int) :: precision, status, s2
  type(c_ptr) :: ctx
  real(c_double), allocatable :: co2v(:,:,:,:,:), co2i(:,:,:,:,:)
  real(c_double) :: pfgasref(MGCM_RT_NP), tgasref(MGCM_RT_NT), wrefh2o(MGCM_RT_NH)
  real(c_double) :: gwv(MGCM_RT_NG, MGCM_RT_NV), gwi(MGCM_RT_NG, MGCM_RT_NI)
  real(c_double) :: fzerov(MGCM_RT_NV), fzeroi(MGCM_RT_NI), tauray(MGCM_RT_NV), dwni(MGCM_RT_NI)
  real(c_double), allocatable :: planckir(:,:)
  real(c_double) :: qxvdst(MGCM_RT_NV), qsvdst(MGCM_RT_NV), gvdst(MGCM_RT_NV)
  real(c_double) :: qxidst(MGCM_RT_NI), qsidst(MGCM_RT_NI), gidst(MGCM_RT_NI)
  real(c_double) :: qxvcld(MGCM_RT_NV), qsvcld(MGCM_RT_NV), gvcld(MGCM_RT_NV)
  real(c_double) :: qxicld(MGCM_RT_NI), qsicld(MGCM_RT_NI), gicld(MGCM_RT_NI)
  real(c_double) :: qextrefdst, qextrefcld, sol(MGCM_RT_NV)
  real(c_double), allocatable :: plev(:,:), tlev(:,:), pmid(:,:), tmid(:,:), qh2o(:,:)
  real(c_double), allocatable :: tauref(:,:), taurefcld(:,:), als(:), albi(:), acosz(:)
  real(c_double), allocatable :: fmnetv(:,:), fluxupv(:,:), fluxdnv(:,:), fmneti(:,:), fluxupi(:,:)
  real(c_double), allocatable :: fluxdni(:,:), nfluxtopv(:), nfluxtopi(:), diffvt(:)
  real(c_double), allocatable :: detau(:,:,:), cumtauv(:,:,:), hr(:,:)
  real(4), allocatable :: fuptopv(:,:), fdntopv(:,:), fupsurfv(:,:), fdnsurfv(:,:)
  real(4), allocatable :: fuptopir(:,:), fupsurfir(:,:), fdnsurfir(:,:)
  integer :: n, a0, nargs
  logical :: percol, aeron, duston
  integer(c_int) :: tile
  character(len=16) :: targ
  character(len=4096) :: ftracer, aerdir, fdust
  real(c_double) :: psf, ptrop, rptau
  real(c_double), allocatable :: tautot(:), conr(:), psurf(:), sigma(:)
  real(c_double), allocatable, target :: qtrace(:,:,:)
  real(c_double), allocatable, target :: dqxv(:,:,:), dqsv(:,:,:), dgv(:,:,:), dqxi(:,:,:), dqsi(:,:,:), dgi(:,:,:)
  real(c_double), allocatable, target :: cqxv(:,:,:), cqsv(:,:,:), cgv(:,:,:), cqxi(:,:,:), cqsi(:,:,:), cgi(:,:,:)
  real(c_double), allocatable, target :: dqref(:,:), cqref(:,:)
  type(mgcm_rt_aerosol), target :: aer
  real(c_double), allocatable :: detau4(:,:,:,:), cumtauv4(:,:,:,:)

  nargs = command_argument_count()
  a0 = 0
  percol = .false.
  aeron = .false.
  duston = .false.
  tile = 0
  do while (a0 < nargs)
    call get_command_argument(a0 + 1, fin)
    if (trim(fin) == '--per-column') then
      percol = .true.
      a0 = a0 + 1
    else if (trim(fin) == '--aerosol' .and. a0 + 3 <= nargs) then
      aeron = .true.
      call get_command_argument(a0 + 2, ftracer)
      call get_command_argument(a0 + 3, aerdir)
      a0 = a0 + 3
    else if (trim(fin) == '--dust' .and. a0 + 2 <= nargs) then
      duston = .true.
      call get_command_argument(a0 + 2, fdust)
      a0 = a0 + 2
    else if (trim(fin) == '--tile' .and. a0 + 2 <= nargs) then
      call get_command_argument(a0 + 2, targ)
      read(targ, *) tile
      a0 = a0 + 2
    else
      exit
    end if
  end do
  if (nargs - a0 < 2 .or. nargs - a0 > 4) then
    write(*, '(a)') 'usage: mgcm_rt_driver [--per-column] [--aerosol <tracer file> <aerosol data dir>] '// &
                    '[--dust <dust file>] [--tile <band slots>] <case file> <result file> '// &
                    '[data directory | -] [history file]'
    stop 2
  end if
  call get_command_argument(a0 + 1, fin)
  call get_command_argument(a0 + 2, fout)
  datadir = ' '
  fhist = ' '
  if (nargs - a0 >= 3) call get_command_argument(a0 + 3, datadir)
  if (trim(datadir) == '-') datadir = ' '
  if (nargs - a0 == 4) call get_command_argument(a0 + 4, fhist)

  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', iostat=ios)
  if (ios /= 0) then
    write(*, '(a)') 'mgcm_rt_driver: cannot open '//trim(fin)
    stop 1
  end if
  read(u) ncol, nlay, precision
  nlev = 2 * nlay + 3
  nl = nlay + 1
  allocate(co2v(MGCM_RT_NT, MGCM_RT_NP, MGCM_RT_NH, MGCM_RT_NV, MGCM_RT_NG))
  allocate(co2i(MGCM_RT_NT, MGCM_RT_NP, MGCM_RT_NH, MGCM_RT_NI, MGCM_RT_NG))
  allocate(planckir(MGCM_RT_NI, MGCM_RT_NPLANCK))
  read(u) co2v, co2i, pfgasref, tgasref, wrefh2o, gwv, gwi, fzerov, fzeroi, tauray, dwni, planckir
  read(u) qxvdst, qsvdst, gvdst, qxidst, qsidst, gidst, qxvcld, qsvcld, gvcld, qxicld, qsicld, gicld
  read(u) qextrefdst, qextrefcld, sol
  allocate(plev(nlev, ncol), tlev(nlev, ncol), pmid(nlev, ncol), tmid(nlev, ncol), qh2o(nlev, ncol))
  allocate(tauref(nlev + 1, ncol), taurefcld(nlev + 1, ncol), als(ncol), albi(ncol), acosz(ncol))
  read(u) plev, tlev, pmid, tmid, qh2o, tauref, taurefcld, als, albi, acosz
  close(u)

  allocate(fmnetv(nl, ncol), fluxupv(nl, ncol), fluxdnv(nl, ncol), fmneti(nl, ncol), fluxupi(nl, ncol))
  allocate(fluxdni(nl, ncol), nfluxtopv(ncol), nfluxtopi(ncol), diffvt(ncol), hr(nl, ncol))
  allocate(detau(MGCM_RT_NV, MGCM_RT_NG, ncol), cumtauv(MGCM_RT_NV, MGCM_RT_NG, ncol))
  detau = 0.0d0
  cumtauv = 0.0d0

  status = MGCM_RT_OK
  if (duston) then
    ! COMP3 without active dust: the prescribed dust's TAUREF (code/comp3.f:591-593)
    allocate(tautot(ncol), conr(ncol), psurf(ncol), sigma(nlev))
    open(newunit=u, file=trim(fdust), access='stream', form='unformatted', status='old', iostat=ios)
    if (ios /= 0) then
      status = MGCM_RT_ERR_IO
    else
      read(u) psf, ptrop, rptau, tautot, conr, psurf, sigma
      close(u)
      call mgcm_rt_dust_profile(ncol, nlay, psf, ptrop, rptau, 100, tautot, conr, psurf, sigma, tauref, status)
    end if
  end if
  if (status == MGCM_RT_OK .and. len_trim(datadir) > 0) call mgcm_rt_laginterp_files(datadir, pgref, pfgasref, co2i, co2v, fzeroi, fzerov, status)
  if (status == MGCM_RT_OK) &
    call mgcm_rt_setup(co2v, co2i, pfgasref, tgasref, wrefh2o, gwv, gwi, fzerov, fzeroi, tauray, dwni, &
                       planckir, qxvdst, qsvdst, gvdst, qxidst, qsidst, gidst, qxvcld, qsvcld, gvcld, &
                       qxicld, qsicld, gicld, qextrefdst, qextrefcld, ctx, status)
  if (status == MGCM_RT_OK) status = mgcm_rt_set_tile(ctx, tile)
  if (status == MGCM_RT_OK .and. aeron) then
    ! COMP3's active dust and clouds: OPT_DST, OPT_CLD, then their outputs into the radiation
    allocate(qtrace(nlay, 6, ncol))
    open(newunit=u, file=trim(ftracer), access='stream', form='unformatted', status='old', iostat=ios)
    if (ios /= 0) then
      status = MGCM_RT_ERR_IO
    else
      read(u) qtrace
      close(u)
      call mgcm_rt_aerosol_setup(ctx, aerdir, status)
    end if
    if (status == MGCM_RT_OK) then
      allocate(dqxv(nlev + 1, MGCM_RT_NV, ncol), dqsv(nlev + 1, MGCM_RT_NV, ncol), dgv(nlev + 1, MGCM_RT_NV, ncol))
      allocate(dqxi(nlev + 1, MGCM_RT_NI, ncol), dqsi(nlev + 1, MGCM_RT_NI, ncol), dgi(nlev + 1, MGCM_RT_NI, ncol))
      allocate(cqxv(nlev + 1, MGCM_RT_NV, ncol), cqsv(nlev + 1, MGCM_RT_NV, ncol), cgv(nlev + 1, MGCM_RT_NV, ncol))
      allocate(cqxi(nlev + 1, MGCM_RT_NI, ncol), cqsi(nlev + 1, MGCM_RT_NI, ncol), cgi(nlev + 1, MGCM_RT_NI, ncol))
      allocate(dqref(nlev + 1, ncol), cqref(nlev + 1, ncol))
      call mgcm_rt_opt_aer(ctx, .false., ncol, nlay, plev, qtrace, dqxv, dqsv, dgv, dqxi, dqsi, dgi, dqref, &
                           tauref, status)
    end if
    if (status == MGCM_RT_OK) &
      call mgcm_rt_opt_aer(ctx, .true., ncol, nlay, plev, qtrace, cqxv, cqsv, cgv, cqxi, cqsi, cgi, cqref, &
                           taurefcld, status)
    if (status == MGCM_RT_OK) then
      aer%qxv_dst = c_loc(dqxv)
      aer%qsv_dst = c_loc(dqsv)
      aer%gv_dst = c_loc(dgv)
      aer%qxi_dst = c_loc(dqxi)
      aer%qsi_dst = c_loc(dqsi)
      aer%gi_dst = c_loc(dgi)
      aer%qxv_cld = c_loc(cqxv)
      aer%qsv_cld = c_loc(cqsv)
      aer%gv_cld = c_loc(cgv)
      aer%qxi_cld = c_loc(cqxi)
      aer%qsi_cld = c_loc(cqsi)
      aer%gi_cld = c_loc(cgi)
      aer%qextref_dst = c_loc(dqref)
      aer%qextref_cld = c_loc(cqref)
      call mgcm_rt_radiation(ctx, ncol, nlay, plev, tlev, pmid, tmid, qh2o, tauref, taurefcld, als, albi, &
                             acosz, sol, fmnetv, fluxupv, fluxdnv, fmneti, fluxupi, fluxdni, nfluxtopv, &
                             nfluxtopi, diffvt, detau, cumtauv, hr, precision, status, aerosol=aer)
    end if
    s2 = mgcm_rt_destroy(ctx)
  else if (status == MGCM_RT_OK .and. percol) then
    ! COMP3's loop over (J, I), here J = column, I = 1 (code/comp3.f:411-415,637-738)
    allocate(detau4(ncol, 1, MGCM_RT_NV, MGCM_RT_NG), cumtauv4(ncol, 1, MGCM_RT_NV, MGCM_RT_NG))
    detau4 = 0.0d0
    cumtauv4 = 0.0d0
    hr = 0.0d0
    do n = 1, ncol
      call mgcm_rt_column(ctx, nlay, plev(:, n), tlev(:, n), pmid(:, n), tmid(:, n), qh2o(:, n), tauref(:, n), &
                          taurefcld(:, n), als(n), albi(n), acosz(n), sol, detau4, cumtauv4, n, 1, &
                          nfluxtopv(n), fmnetv(:, n), fluxupv(:, n), fluxdnv(:, n), diffvt(n), &
                          nfluxtopi(n), fmneti(:, n), fluxupi(:, n), fluxdni(:, n), status)
      if (status /= MGCM_RT_OK) exit
    end do
    do n = 1, ncol
      detau(:, :, n) = detau4(n, 1, :, :)
      cumtauv(:, :, n) = cumtauv4(n, 1, :, :)
    end do
    s2 = mgcm_rt_destroy(ctx)
  else if (status == MGCM_RT_OK) then
    call mgcm_rt_radiation(ctx, ncol, nlay, plev, tlev, pmid, tmid, qh2o, tauref, taurefcld, als, albi, &
                           acosz, sol, fmnetv, fluxupv, fluxdnv, fmneti, fluxupi, fluxdni, nfluxtopv, &
                           nfluxtopi, diffvt, detau, cumtauv, hr, precision, status)
    s2 = mgcm_rt_destroy(ctx)
  end if
  if (status /= MGCM_RT_OK) write(*, '(a,i0,a)') 'mgcm_rt_driver: status ', status, ': '//trim(mgcm_rt_message(status))

  if (status == MGCM_RT_OK .and. len_trim(fhist) > 0) then
    allocate(fuptopv(ncol, 1), fdntopv(ncol, 1), fupsurfv(ncol, 1), fdnsurfv(ncol, 1))
    allocate(fuptopir(ncol, 1), fupsurfir(ncol, 1), fdnsurfir(ncol, 1))
    do n = 1, ncol
      if (acosz(n) >= 1.0e-5) then
        fuptopv(n, 1) = fluxupv(1, n)
        fdntopv(n, 1) = fluxdnv(1, n)
        fupsurfv(n, 1) = fluxupv(nl, n)
        fdnsurfv(n, 1) = fluxdnv(nl, n)
      else
        fuptopv(n, 1) = 0.0
        fdntopv(n, 1) = 0.0
        fupsurfv(n, 1) = 0.0
        fdnsurfv(n, 1) = 0.0
      end if
      fuptopir(n, 1) = fluxupi(1, n)
      fupsurfir(n, 1) = fluxupi(nl, n)
      fdnsurfir(n, 1) = fluxdni(nl, n)
    end do
    open(newunit=u, file=trim(fhist), form='unformatted', status='replace')
    write(u) fuptopv, fdntopv, fupsurfv, fdnsurfv
    write(u) fuptopir, fupsurfir, fdnsurfir
    close(u)
  end if

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) status
  write(u) fmnetv, fluxupv, fluxdnv, fmneti, fluxupi, fluxdni, nfluxtopv, nfluxtopi, diffvt, detau, cumtauv, hr
  close(u)
  if (status /= MGCM_RT_OK) stop 1
end program mgcm_rt_driver
