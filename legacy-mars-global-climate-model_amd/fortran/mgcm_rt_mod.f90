! mgcm_rt_mod.f90 -- Fortran binding of the MI355X column radiative transfer
! (include/mgcm_rt.h) through ISO_C_BINDING.
!
! The reference's physics driver COMP3 calls OPTCV, SFLUXV, OPTCI and SFLUXI
! once per column (code/comp3.f:637-738).  With this module it gathers a batch
! of columns in the reference's per-column layouts, calls mgcm_rt_radiation
! once, and scatters the results: FMNETV/FLUXUPV/FLUXDNV, FMNETI/FLUXUPI/
! FLUXDNI per radiation layer, NFLUXTOPV/I, DIFFVT, DETAU and TAUCUMV(L_LEVELS)
! per (band, Gauss point), and the layer heating rates in K/sol.
module mgcm_rt_mod
  use iso_c_binding
  implicit none
  private

  integer, parameter, public :: MGCM_RT_NT = 7, MGCM_RT_NP = 51, MGCM_RT_NH = 10, MGCM_RT_NG = 17
  integer, parameter, public :: MGCM_RT_NV = 7, MGCM_RT_NI = 5, MGCM_RT_NPLANCK = 8501, MGCM_RT_MAX_NLAY = 99
  integer(c_int), parameter, public :: MGCM_RT_OK = 0, MGCM_RT_ERR_ARG = -1, MGCM_RT_ERR_NLAY = -2
  integer(c_int), parameter, public :: MGCM_RT_ERR_DEVICE = -3, MGCM_RT_ERR_NOMEM = -4
  integer(c_int), parameter, public :: MGCM_RT_ERR_RANGE = -5, MGCM_RT_ERR_PRECISION = -6, MGCM_RT_ERR_IO = -7
  integer(c_int), parameter, public :: MGCM_RT_FP64 = 0, MGCM_RT_MIXED = 1
  ! added to the precision: run one part only (the other part's outputs are not written)
  integer(c_int), parameter, public :: MGCM_RT_SKIP_VIS = 16, MGCM_RT_SKIP_IR = 32

  ! mirrors mgcm_rt_tables (host pointers, copied by mgcm_rt_init)
  type, bind(C), public :: mgcm_rt_tables
    type(c_ptr) :: co2v, co2i, pfgasref, tgasref, wrefh2o, gwv, gwi, fzerov, fzeroi, tauray, dwni, planckir
    type(c_ptr) :: qxv_dst, qsv_dst, gv_dst, qxi_dst, qsi_dst, gi_dst
    type(c_ptr) :: qxv_cld, qsv_cld, gv_cld, qxi_cld, qsi_cld, gi_cld
    real(c_double) :: qextref_dst, qextref_cld
  end type

  ! mirrors mgcm_rt_aerosol: per-level aerosol optics of each column (OPT_DST/OPT_CLD
  ! outputs, Q(L_LEVELS+1, nspect) per column); c_null_ptr = the tables' constants
  type, bind(C), public :: mgcm_rt_aerosol
    type(c_ptr) :: qxv_dst, qsv_dst, gv_dst, qxi_dst, qsi_dst, gi_dst
    type(c_ptr) :: qxv_cld, qsv_cld, gv_cld, qxi_cld, qsi_cld, gi_cld
    type(c_ptr) :: qextref_dst, qextref_cld
  end type

  ! mirrors mgcm_rt_columns
  type, bind(C), public :: mgcm_rt_columns
    integer(c_int) :: ncol, nlay
    type(c_ptr) :: plev, tlev, pmid, tmid, qh2o, tauref, taurefcld, als, albi, acosz
    real(c_double) :: sol(7)
    type(c_ptr) :: aerosol
  end type

  ! mirrors mgcm_rt_fluxes
  type, bind(C), public :: mgcm_rt_fluxes
    type(c_ptr) :: fmnetv, fluxupv, fluxdnv, fmneti, fluxupi, fluxdni
    type(c_ptr) :: nfluxtopv, nfluxtopi, diffvt, detau, cumtauv, hr
    type(c_ptr) :: directsol, dnvflux
  end type

  ! mirrors mgcm_rt_history: COMP3's REAL*4 history fields and the non-LTE
  ! heating of a device-resident run (mgcm_rt_history_fields)
  type, bind(C), public :: mgcm_rt_history
    type(c_ptr) :: fuptopv, fdntopv, fupsurfv, fdnsurfv, fuptopir, fupsurfir, fdnsurfir, surfalb
    type(c_ptr) :: hr_nlte, p, sigma
    real(c_double) :: ptrop
  end type

  public :: mgcm_rt_init, mgcm_rt_destroy, mgcm_rt_run, mgcm_rt_run_host, mgcm_rt_check, mgcm_rt_history_fields
  public :: mgcm_rt_run_batched, mgcm_rt_planck_table, mgcm_rt_laginterp, mgcm_rt_read_ktable, mgcm_rt_strerror
  public :: mgcm_rt_set_tile, mgcm_rt_dsolflux_host
  ! mirrors mgcm_rt_tracers: QTRACE(nlay, 6) and PL(L_LEVELS) of each column
  type, bind(C), public :: mgcm_rt_tracers
    integer(c_int) :: ncol, nlay
    type(c_ptr) :: pl, qtrace
  end type

  ! mirrors mgcm_rt_aerosol_out: OPT_DST / OPT_CLD outputs of each column
  type, bind(C), public :: mgcm_rt_aerosol_out
    type(c_ptr) :: qxv, qsv, gv, qxi, qsi, gi, qextref, tauref, tau2
  end type

  ! mirrors mgcm_rt_dust_columns: the prescribed (Conrath) dust of each column
  ! (DUSTPROFILE + FILLTAUCUM, code/dustprofile.f, code/filltaucum.f)
  type, bind(C), public :: mgcm_rt_dust_columns
    integer(c_int) :: ncol, nlay
    type(c_ptr) :: tautot, conr, p, sigma, prdst
    integer(c_int) :: npdst
    real(c_double) :: rptau, ptrop
  end type

  integer, parameter, public :: MGCM_RT_AER_TABLE_DOUBLES = 11552

  public :: mgcm_rt_setup, mgcm_rt_radiation, mgcm_rt_message, mgcm_rt_laginterp_files, mgcm_rt_column
  public :: mgcm_rt_read_aerosol_tables, mgcm_rt_set_aerosol_tables, mgcm_rt_opt_dst_host, mgcm_rt_opt_cld_host
  public :: mgcm_rt_aerosol_setup, mgcm_rt_opt_aer
  public :: mgcm_rt_dust_grid, mgcm_rt_dust_tauref_host, mgcm_rt_dust_profile

  interface
    integer(c_int) function mgcm_rt_init(tables, ctx) bind(C, name="mgcm_rt_init")
      import :: c_int, c_ptr, mgcm_rt_tables
      type(mgcm_rt_tables), intent(in) :: tables
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function mgcm_rt_destroy(ctx) bind(C, name="mgcm_rt_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mgcm_rt_run(ctx, cols, outs, precision, stream) bind(C, name="mgcm_rt_run")
      import :: c_int, c_ptr, mgcm_rt_columns, mgcm_rt_fluxes
      type(c_ptr), value :: ctx
      type(mgcm_rt_columns), intent(in) :: cols
      type(mgcm_rt_fluxes), intent(in) :: outs
      integer(c_int), value :: precision
      type(c_ptr), value :: stream
    end function
    integer(c_int) function mgcm_rt_run_host(ctx, cols, outs, precision) bind(C, name="mgcm_rt_run_host")
      import :: c_int, c_ptr, mgcm_rt_columns, mgcm_rt_fluxes
      type(c_ptr), value :: ctx
      type(mgcm_rt_columns), intent(in) :: cols
      type(mgcm_rt_fluxes), intent(in) :: outs
      integer(c_int), value :: precision
    end function
    integer(c_int) function mgcm_rt_check(ctx) bind(C, name="mgcm_rt_check")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    ! DSOLFLUX + DNVFLUX = DIFFVT + DIRECTSOL of ncol columns from host arrays
    ! (code/comp3.f:311-351): detau(7, 17, ncol) as DETAU(J,I,:,:) of each column
    integer(c_int) function mgcm_rt_dsolflux_host(ctx, ncol, sol, acosz, detau, diffvt, directsol, dnvflux) &
        bind(C, name="mgcm_rt_dsolflux_host")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ncol
      real(c_double), intent(in) :: sol(*), acosz(*), detau(*), diffvt(*)
      real(c_double), intent(out) :: directsol(*), dnvflux(*)
    end function
    ! column kernel tile of single-grid calls: 1, 2 or 4 band slots per column, 12 = one band per wave,
    ! 0 = automatic
    integer(c_int) function mgcm_rt_set_tile(ctx, band_slots) bind(C, name="mgcm_rt_set_tile")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: band_slots
    end function
    integer(c_int) function mgcm_rt_run_batched(ctx, ngrids, grids, outs, precision, stream) &
        bind(C, name="mgcm_rt_run_batched")
      import :: c_int, c_ptr, mgcm_rt_columns, mgcm_rt_fluxes
      type(c_ptr), value :: ctx
      integer(c_int), value :: ngrids
      type(mgcm_rt_columns), intent(in) :: grids(*)
      type(mgcm_rt_fluxes), intent(in) :: outs(*)
      integer(c_int), value :: precision
      type(c_ptr), value :: stream
    end function
    integer(c_int) function mgcm_rt_planck_table(planckir, dwni, wnoi) bind(C, name="mgcm_rt_planck_table")
      import :: c_int, c_double
      real(c_double), intent(out) :: planckir(*), dwni(*), wnoi(*)
    end function
    integer(c_int) function mgcm_rt_laginterp(pgref, raw, nbands, pint, out) bind(C, name="mgcm_rt_laginterp")
      import :: c_int, c_double
      real(c_double), intent(in) :: pgref(*), raw(*)
      integer(c_int), value :: nbands
      real(c_double), intent(out) :: pint(*), out(*)
    end function
    integer(c_int) function mgcm_rt_read_ktable(path, nbands, raw, fzero) bind(C, name="mgcm_rt_read_ktable")
      import :: c_int, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: nbands
      real(c_double), intent(out) :: raw(*), fzero(*)
    end function
    integer(c_int) function mgcm_rt_history_fields(cols, outs, hist, stream) bind(C, name="mgcm_rt_history_fields")
      import :: c_int, c_ptr, mgcm_rt_columns, mgcm_rt_fluxes, mgcm_rt_history
      type(mgcm_rt_columns), intent(in) :: cols
      type(mgcm_rt_fluxes), intent(in) :: outs
      type(mgcm_rt_history), intent(in) :: hist
      type(c_ptr), value :: stream
    end function
    integer(c_int) function mgcm_rt_read_aerosol_tables(datadir, tables) bind(C, name="mgcm_rt_read_aerosol_tables")
      import :: c_int, c_double, c_char
      character(kind=c_char), intent(in) :: datadir(*)
      real(c_double), intent(out) :: tables(*)
    end function
    integer(c_int) function mgcm_rt_set_aerosol_tables(ctx, tables) bind(C, name="mgcm_rt_set_aerosol_tables")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: tables(*)
    end function
    integer(c_int) function mgcm_rt_opt_dst_host(ctx, tr, out) bind(C, name="mgcm_rt_opt_dst_host")
      import :: c_int, c_ptr, mgcm_rt_tracers, mgcm_rt_aerosol_out
      type(c_ptr), value :: ctx
      type(mgcm_rt_tracers), intent(in) :: tr
      type(mgcm_rt_aerosol_out), intent(in) :: out
    end function
    integer(c_int) function mgcm_rt_opt_cld_host(ctx, tr, out) bind(C, name="mgcm_rt_opt_cld_host")
      import :: c_int, c_ptr, mgcm_rt_tracers, mgcm_rt_aerosol_out
      type(c_ptr), value :: ctx
      type(mgcm_rt_tracers), intent(in) :: tr
      type(mgcm_rt_aerosol_out), intent(in) :: out
    end function
    integer(c_int) function mgcm_rt_dust_grid(psf, ptrop, npdst, prdst) bind(C, name="mgcm_rt_dust_grid")
      import :: c_int, c_double
      real(c_double), value :: psf, ptrop
      integer(c_int), value :: npdst
      real(c_double), intent(out) :: prdst(*)
    end function
    integer(c_int) function mgcm_rt_dust_tauref_host(dc, tauref, taudst) bind(C, name="mgcm_rt_dust_tauref_host")
      import :: c_int, c_ptr, mgcm_rt_dust_columns
      type(mgcm_rt_dust_columns), intent(in) :: dc
      type(c_ptr), value :: tauref, taudst
    end function
    type(c_ptr) function mgcm_rt_strerror(status) bind(C, name="mgcm_rt_strerror")
      import :: c_int, c_ptr
      integer(c_int), value :: status
    end function
  end interface

contains

  ! Status text of a return code.
  function mgcm_rt_message(status) result(msg)
    integer(c_int), intent(in) :: status
    character(len=80) :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ' '
    call c_f_pointer(mgcm_rt_strerror(status), s, [80])
    do i = 1, 80
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function

  ! LAGINTERP (code/laginterp.f90:1-239) with its own arguments, from the data
  ! directory that holds CO2H2O_V_12_95_INTEL and CO2H2O_IR_12_95_INTEL: the
  ! library reads the files and refines the tables from the 11 reference
  ! pressures pgref (mbar) to the 51 PINT points.  status: MGCM_RT_OK or the
  ! library's error code (MGCM_RT_ERR_IO for a missing or malformed file).
  subroutine mgcm_rt_laginterp_files(datadir, pgref, pint, co2i, co2v, fzeroi, fzerov, status)
    character(len=*), intent(in) :: datadir
    real(c_double), intent(in) :: pgref(11)
    real(c_double), intent(out) :: pint(MGCM_RT_NP)
    real(c_double), intent(out) :: co2i(MGCM_RT_NT, MGCM_RT_NP, MGCM_RT_NH, MGCM_RT_NI, MGCM_RT_NG)
    real(c_double), intent(out) :: co2v(MGCM_RT_NT, MGCM_RT_NP, MGCM_RT_NH, MGCM_RT_NV, MGCM_RT_NG)
    real(c_double), intent(out) :: fzeroi(MGCM_RT_NI), fzerov(MGCM_RT_NV)
    integer(c_int), intent(out) :: status
    real(c_double), allocatable :: raw(:)
    allocate(raw(MGCM_RT_NT * 11 * MGCM_RT_NH * MGCM_RT_NV * MGCM_RT_NG))
    status = mgcm_rt_read_ktable(trim(datadir)//'/CO2H2O_V_12_95_INTEL'//c_null_char, int(MGCM_RT_NV, c_int), &
                                 raw, fzerov)
    if (status == MGCM_RT_OK) status = mgcm_rt_laginterp(pgref, raw, int(MGCM_RT_NV, c_int), pint, co2v)
    if (status == MGCM_RT_OK) &
      status = mgcm_rt_read_ktable(trim(datadir)//'/CO2H2O_IR_12_95_INTEL'//c_null_char, int(MGCM_RT_NI, c_int), &
                                   raw, fzeroi)
    if (status == MGCM_RT_OK) status = mgcm_rt_laginterp(pgref, raw, int(MGCM_RT_NI, c_int), pint, co2i)
    deallocate(raw)
  end subroutine

  ! RADSETUP's tables (code/radsetup.f90:26-46) into a device context.  Arrays
  ! in the reference's layouts; the aerosol constants per band.
  subroutine mgcm_rt_setup(co2v, co2i, pfgasref, tgasref, wrefh2o, gwv, gwi, fzerov, fzeroi, tauray, dwni, &
                           planckir, qxvdst, qsvdst, gvdst, qxidst, qsidst, gidst, qxvcld, qsvcld, gvcld, &
                           qxicld, qsicld, gicld, qextrefdst, qextrefcld, ctx, status)
    real(c_double), intent(in), target :: co2v(MGCM_RT_NT, MGCM_RT_NP, MGCM_RT_NH, MGCM_RT_NV, MGCM_RT_NG)
    real(c_double), intent(in), target :: co2i(MGCM_RT_NT, MGCM_RT_NP, MGCM_RT_NH, MGCM_RT_NI, MGCM_RT_NG)
    real(c_double), intent(in), target :: pfgasref(MGCM_RT_NP), tgasref(MGCM_RT_NT), wrefh2o(MGCM_RT_NH)
    real(c_double), intent(in), target :: gwv(MGCM_RT_NG, MGCM_RT_NV), gwi(MGCM_RT_NG, MGCM_RT_NI)
    real(c_double), intent(in), target :: fzerov(MGCM_RT_NV), fzeroi(MGCM_RT_NI), tauray(MGCM_RT_NV)
    real(c_double), intent(in), target :: dwni(MGCM_RT_NI), planckir(MGCM_RT_NI, MGCM_RT_NPLANCK)
    real(c_double), intent(in), target :: qxvdst(MGCM_RT_NV), qsvdst(MGCM_RT_NV), gvdst(MGCM_RT_NV)
    real(c_double), intent(in), target :: qxidst(MGCM_RT_NI), qsidst(MGCM_RT_NI), gidst(MGCM_RT_NI)
    real(c_double), intent(in), target :: qxvcld(MGCM_RT_NV), qsvcld(MGCM_RT_NV), gvcld(MGCM_RT_NV)
    real(c_double), intent(in), target :: qxicld(MGCM_RT_NI), qsicld(MGCM_RT_NI), gicld(MGCM_RT_NI)
    real(c_double), intent(in) :: qextrefdst, qextrefcld
    type(c_ptr), intent(out) :: ctx
    integer(c_int), intent(out) :: status
    type(mgcm_rt_tables) :: t
    t%co2v = c_loc(co2v)
    t%co2i = c_loc(co2i)
    t%pfgasref = c_loc(pfgasref)
    t%tgasref = c_loc(tgasref)
    t%wrefh2o = c_loc(wrefh2o)
    t%gwv = c_loc(gwv)
    t%gwi = c_loc(gwi)
    t%fzerov = c_loc(fzerov)
    t%fzeroi = c_loc(fzeroi)
    t%tauray = c_loc(tauray)
    t%dwni = c_loc(dwni)
    t%planckir = c_loc(planckir)
    t%qxv_dst = c_loc(qxvdst)
    t%qsv_dst = c_loc(qsvdst)
    t%gv_dst = c_loc(gvdst)
    t%qxi_dst = c_loc(qxidst)
    t%qsi_dst = c_loc(qsidst)
    t%gi_dst = c_loc(gidst)
    t%qxv_cld = c_loc(qxvcld)
    t%qsv_cld = c_loc(qsvcld)
    t%gv_cld = c_loc(gvcld)
    t%qxi_cld = c_loc(qxicld)
    t%qsi_cld = c_loc(qsicld)
    t%gi_cld = c_loc(gicld)
    t%qextref_dst = qextrefdst
    t%qextref_cld = qextrefcld
    status = mgcm_rt_init(t, ctx)
  end subroutine

  ! The per-column OPTCV/SFLUXV/OPTCI/SFLUXI sequence of COMP3 for ncol
  ! columns at once (host arrays; copies in, runs, copies out).  Column i
  ! holds PLEV(:,i) etc. as COMP3 passes them; SOL = SOLARF/RSDIST.
  ! detau/cumtauv keep their values for sun-down columns, as the reference
  ! leaves its module arrays.
  subroutine mgcm_rt_radiation(ctx, ncol, nlay, plev, tlev, pmid, tmid, qh2o, tauref, taurefcld, &
                               als, albi, acosz, sol, fmnetv, fluxupv, fluxdnv, fmneti, fluxupi, &
                               fluxdni, nfluxtopv, nfluxtopi, diffvt, detau, cumtauv, hr, precision, status, &
                               aerosol, directsol, dnvflux)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: ncol, nlay
    real(c_double), intent(in), target :: plev(2*nlay+3, ncol), tlev(2*nlay+3, ncol)
    real(c_double), intent(in), target :: pmid(2*nlay+3, ncol), tmid(2*nlay+3, ncol), qh2o(2*nlay+3, ncol)
    real(c_double), intent(in), target :: tauref(2*nlay+4, ncol), taurefcld(2*nlay+4, ncol)
    real(c_double), intent(in), target :: als(ncol), albi(ncol), acosz(ncol)
    real(c_double), intent(in) :: sol(MGCM_RT_NV)
    real(c_double), intent(out), target :: fmnetv(nlay+1, ncol), fluxupv(nlay+1, ncol), fluxdnv(nlay+1, ncol)
    real(c_double), intent(out), target :: fmneti(nlay+1, ncol), fluxupi(nlay+1, ncol), fluxdni(nlay+1, ncol)
    real(c_double), intent(out), target :: nfluxtopv(ncol), nfluxtopi(ncol), diffvt(ncol)
    real(c_double), intent(inout), target :: detau(MGCM_RT_NV, MGCM_RT_NG, ncol)
    real(c_double), intent(inout), target :: cumtauv(MGCM_RT_NV, MGCM_RT_NG, ncol)
    real(c_double), intent(out), target :: hr(nlay+1, ncol)
    integer(c_int), intent(in) :: precision
    integer(c_int), intent(out) :: status
    ! per-level aerosol optics (active dust/clouds); absent = the tables' constants
    type(mgcm_rt_aerosol), intent(in), target, optional :: aerosol
    ! DSOLFLUX's direct surface solar flux and DNVFLUX (code/comp3.f:336-348), optional
    real(c_double), intent(out), target, optional :: directsol(ncol), dnvflux(ncol)
    type(mgcm_rt_columns) :: c
    type(mgcm_rt_fluxes) :: o
    c%ncol = ncol
    c%nlay = nlay
    c%plev = c_loc(plev)
    c%tlev = c_loc(tlev)
    c%pmid = c_loc(pmid)
    c%tmid = c_loc(tmid)
    c%qh2o = c_loc(qh2o)
    c%tauref = c_loc(tauref)
    c%taurefcld = c_loc(taurefcld)
    c%als = c_loc(als)
    c%albi = c_loc(albi)
    c%acosz = c_loc(acosz)
    c%sol = sol
    c%aerosol = c_null_ptr
    if (present(aerosol)) c%aerosol = c_loc(aerosol)
    o%fmnetv = c_loc(fmnetv)
    o%fluxupv = c_loc(fluxupv)
    o%fluxdnv = c_loc(fluxdnv)
    o%fmneti = c_loc(fmneti)
    o%fluxupi = c_loc(fluxupi)
    o%fluxdni = c_loc(fluxdni)
    o%nfluxtopv = c_loc(nfluxtopv)
    o%nfluxtopi = c_loc(nfluxtopi)
    o%diffvt = c_loc(diffvt)
    o%detau = c_loc(detau)
    o%cumtauv = c_loc(cumtauv)
    o%hr = c_loc(hr)
    o%directsol = c_null_ptr
    o%dnvflux = c_null_ptr
    if (present(directsol)) o%directsol = c_loc(directsol)
    if (present(directsol) .and. present(dnvflux)) o%dnvflux = c_loc(dnvflux)
    status = mgcm_rt_run_host(ctx, c, o, precision)
  end subroutine

  ! INITCLD's size grid and optical tables (code/initcld.f:131-184) from the four
  ! data files in datadir (Dust_*_wolff2010_JD_12bands.dat,
  ! waterCoated_*_JD_12bands.dat), uploaded to the context.
  subroutine mgcm_rt_aerosol_setup(ctx, datadir, status)
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: datadir
    integer(c_int), intent(out) :: status
    real(c_double), allocatable :: tables(:)
    allocate(tables(MGCM_RT_AER_TABLE_DOUBLES))
    status = mgcm_rt_read_aerosol_tables(trim(datadir)//c_null_char, tables)
    if (status == MGCM_RT_OK) status = mgcm_rt_set_aerosol_tables(ctx, tables)
    deallocate(tables)
  end subroutine

  ! OPT_DST (cloud = .false., code/opt_dst.f) or OPT_CLD (code/opt_cld.f) for a
  ! batch of columns, host arrays: PL(L_LEVELS, ncol) and QTRACE(nlay, 6, ncol) in;
  ! Qext/Qscat/g (L_LEVELS+1, nspect, ncol), the reference extinction and
  ! TAUREF (L_LEVELS+1, ncol) and, optionally, taudust/taucloud (2, ncol) out --
  ! the per-level optics mgcm_rt_radiation takes through its aerosol argument.
  subroutine mgcm_rt_opt_aer(ctx, cloud, ncol, nlay, pl, qtrace, qxv, qsv, gv, qxi, qsi, gi, qextref, tauref, &
                             status, tau2)
    type(c_ptr), intent(in) :: ctx
    logical, intent(in) :: cloud
    integer, intent(in) :: ncol, nlay
    real(c_double), intent(in), target :: pl(2*nlay+3, ncol), qtrace(nlay, 6, ncol)
    real(c_double), intent(out), target :: qxv(2*nlay+4, MGCM_RT_NV, ncol), qsv(2*nlay+4, MGCM_RT_NV, ncol)
    real(c_double), intent(out), target :: gv(2*nlay+4, MGCM_RT_NV, ncol), qxi(2*nlay+4, MGCM_RT_NI, ncol)
    real(c_double), intent(out), target :: qsi(2*nlay+4, MGCM_RT_NI, ncol), gi(2*nlay+4, MGCM_RT_NI, ncol)
    real(c_double), intent(out), target :: qextref(2*nlay+4, ncol), tauref(2*nlay+4, ncol)
    integer(c_int), intent(out) :: status
    real(c_double), intent(out), target, optional :: tau2(2, ncol)
    type(mgcm_rt_tracers) :: tr
    type(mgcm_rt_aerosol_out) :: o
    tr%ncol = ncol
    tr%nlay = nlay
    tr%pl = c_loc(pl)
    tr%qtrace = c_loc(qtrace)
    o%qxv = c_loc(qxv)
    o%qsv = c_loc(qsv)
    o%gv = c_loc(gv)
    o%qxi = c_loc(qxi)
    o%qsi = c_loc(qsi)
    o%gi = c_loc(gi)
    o%qextref = c_loc(qextref)
    o%tauref = c_loc(tauref)
    o%tau2 = c_null_ptr
    if (present(tau2)) o%tau2 = c_loc(tau2)
    if (cloud) then
      status = mgcm_rt_opt_cld_host(ctx, tr, o)
    else
      status = mgcm_rt_opt_dst_host(ctx, tr, o)
    end if
  end subroutine

  ! Prescribed dust as DUSTPROFILE and FILLTAUCUM give it (code/dustprofile.f:
  ! 30-113, code/filltaucum.f:33-52): PRDST(npdst) from PSF and PTROP, then each
  ! column's Conrath profile with opacity TAUTOT at RPTAU and parameter CONR,
  ! surface pressure P = PS - PTROP, on the sub-levels of SIGMA(L_LEVELS):
  ! TAUREF(L_LEVELS+1, ncol) and, optionally, TAUDST(npdst, ncol).
  subroutine mgcm_rt_dust_profile(ncol, nlay, psf, ptrop, rptau, npdst, tautot, conr, p, sigma, tauref, &
                                  status, taudst)
    integer, intent(in) :: ncol, nlay, npdst
    real(c_double), intent(in) :: psf, ptrop, rptau
    real(c_double), intent(in), target :: tautot(ncol), conr(ncol), p(ncol), sigma(2*nlay+3)
    real(c_double), intent(out), target :: tauref(2*nlay+4, ncol)
    integer(c_int), intent(out) :: status
    real(c_double), intent(out), target, optional :: taudst(npdst, ncol)
    real(c_double), allocatable, target :: prdst(:)
    type(mgcm_rt_dust_columns) :: dc
    allocate(prdst(npdst))
    status = mgcm_rt_dust_grid(psf, ptrop, int(npdst, c_int), prdst)
    if (status /= MGCM_RT_OK) return
    dc%ncol = ncol
    dc%nlay = nlay
    dc%tautot = c_loc(tautot)
    dc%conr = c_loc(conr)
    dc%p = c_loc(p)
    dc%sigma = c_loc(sigma)
    dc%prdst = c_loc(prdst)
    dc%npdst = npdst
    dc%rptau = rptau
    dc%ptrop = ptrop
    if (present(taudst)) then
      status = mgcm_rt_dust_tauref_host(dc, c_loc(tauref), c_loc(taudst))
    else
      status = mgcm_rt_dust_tauref_host(dc, c_loc(tauref), c_null_ptr)
    end if
  end subroutine

  ! Per-column drop-in for COMP3's radiation block (code/comp3.f:637-738): one
  ! call in place of
  !     if (ACOSZ >= 1.0e-5) then; call optcv(...); call sfluxv(...); end if
  !     call optci(...); call sfluxi(...)
  ! for column (J, I), with COMP3's per-column arrays in the reference's shapes:
  ! PLEV, TLEV, PMID, TMID, QH2O(L_LEVELS), TAUREF, TAUREFCLD(L_LEVELS+1), the
  ! fluxes (L_NLAYRAD).  DETAU and CUMTAUV are the module arrays
  ! (L_J, L_I, L_NSPECTV, L_NGAUSS) (code/modules.f90:319,410), written at (J, I)
  ! as SFLUXV writes detau(j1d,i1d,nw,ng) (code/sfluxv.f90:30,68-70) and COMP3
  ! CUMTAUV (code/comp3.f:664-668), and left as they are when the sun is down;
  ! the VIS outputs of a sun-down column are zero.  One device run per column:
  ! for porting COMP3 a step at a time and for drop-in tests; mgcm_rt_radiation
  ! on a whole batch is the fast path.
  subroutine mgcm_rt_column(ctx, nlay, plev, tlev, pmid, tmid, qh2o, tauref, taurefcld, als, albi, acosz, &
                            sol, detau, cumtauv, j, i, nfluxtopv, fmnetv, fluxupv, fluxdnv, diffvt, &
                            nfluxtopi, fmneti, fluxupi, fluxdni, status, aerosol)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nlay, j, i
    real(c_double), intent(in) :: plev(2*nlay+3), tlev(2*nlay+3), pmid(2*nlay+3), tmid(2*nlay+3)
    real(c_double), intent(in) :: qh2o(2*nlay+3), tauref(2*nlay+4), taurefcld(2*nlay+4)
    real(c_double), intent(in) :: als, albi, acosz, sol(MGCM_RT_NV)
    real(c_double), intent(inout) :: detau(:,:,:,:), cumtauv(:,:,:,:)
    real(c_double), intent(out) :: nfluxtopv, fmnetv(nlay+1), fluxupv(nlay+1), fluxdnv(nlay+1), diffvt
    real(c_double), intent(out) :: nfluxtopi, fmneti(nlay+1), fluxupi(nlay+1), fluxdni(nlay+1)
    integer(c_int), intent(out) :: status
    type(mgcm_rt_aerosol), intent(in), optional :: aerosol
    real(c_double) :: pl(2*nlay+3, 1), tl(2*nlay+3, 1), pm(2*nlay+3, 1), tm(2*nlay+3, 1), qh(2*nlay+3, 1)
    real(c_double) :: tr(2*nlay+4, 1), tc(2*nlay+4, 1), a1(1), a2(1), a3(1)
    real(c_double) :: fv(nlay+1, 1), uv(nlay+1, 1), dv(nlay+1, 1), fi(nlay+1, 1), ui(nlay+1, 1), di(nlay+1, 1)
    real(c_double) :: ntv(1), nti(1), dif(1), hr(nlay+1, 1)
    real(c_double) :: det(MGCM_RT_NV, MGCM_RT_NG, 1), cum(MGCM_RT_NV, MGCM_RT_NG, 1)
    integer :: nw, ng
    pl(:, 1) = plev
    tl(:, 1) = tlev
    pm(:, 1) = pmid
    tm(:, 1) = tmid
    qh(:, 1) = qh2o
    tr(:, 1) = tauref
    tc(:, 1) = taurefcld
    a1(1) = als
    a2(1) = albi
    a3(1) = acosz
    do ng = 1, MGCM_RT_NG
      do nw = 1, MGCM_RT_NV
        det(nw, ng, 1) = detau(j, i, nw, ng)
        cum(nw, ng, 1) = cumtauv(j, i, nw, ng)
      end do
    end do
    if (present(aerosol)) then
      call mgcm_rt_radiation(ctx, 1, nlay, pl, tl, pm, tm, qh, tr, tc, a1, a2, a3, sol, fv, uv, dv, fi, ui, di, &
                             ntv, nti, dif, det, cum, hr, MGCM_RT_FP64, status, aerosol=aerosol)
    else
      call mgcm_rt_radiation(ctx, 1, nlay, pl, tl, pm, tm, qh, tr, tc, a1, a2, a3, sol, fv, uv, dv, fi, ui, di, &
                             ntv, nti, dif, det, cum, hr, MGCM_RT_FP64, status)
    end if
    fmnetv = fv(:, 1)
    fluxupv = uv(:, 1)
    fluxdnv = dv(:, 1)
    fmneti = fi(:, 1)
    fluxupi = ui(:, 1)
    fluxdni = di(:, 1)
    nfluxtopv = ntv(1)
    nfluxtopi = nti(1)
    diffvt = dif(1)
    do ng = 1, MGCM_RT_NG
      do nw = 1, MGCM_RT_NV
        detau(j, i, nw, ng) = det(nw, ng, 1)
        cumtauv(j, i, nw, ng) = cum(nw, ng, 1)
      end do
    end do
  end subroutine

end module mgcm_rt_mod
