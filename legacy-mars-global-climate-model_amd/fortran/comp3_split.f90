! comp3_split.f90 -- COMP3's column loop in two passes around one batched
! radiation call, for a build of the GCM itself (oracle/build_model.sh makes
! the loop edit in a temporary copy of code/comp3.f).
!
! The loop body (code/comp3.f:417-1042) splits at the radiation calls:
!   pass 1, every column (I outer, J inner, as COMP3 loops): the part before the
!     radiation (:417-636: winds, potential temperature, CO2 condensation,
!     FILLPT, water, dust and cloud optics) and the two albedos (:642-653 with
!     the sun up, :711-720); mgcm_comp3_record keeps the column's radiation
!     inputs and comp3_split_save the state that part leaves for the rest;
!   comp3_split_radiate: every recorded column in one call -- the library on the
!     MI355X (comp3_rt_batch.f90), or the reference's own routines column by
!     column in the CPU twin (oracle/comp3_ref_split.f90);
!   pass 2, every column in the same order: comp3_split_restore, then the rest
!     of the body (:637-1042) unchanged, whose radiation call sites
!     (mgcm_comp3_vis, mgcm_comp3_ir) return the batch's results for the column
!     after checking, bit for bit, that COMP3 hands them the inputs pass 1
!     recorded.
!
! Saved per column: what the first part writes and the rest reads.  COMP3's
! locals QPI, QPIG, TETAORG, QPISAV and DUSTREF, and the radiation grid PLEV,
! TLEV, PMID, TMID, QH2O; from comp3cmn_h the indices ICMN, IM1, IP1, JCMN,
! JM1, JP1, ACOSZ, PSAT, TSAT, TAUICET, TAUICEL, the winds UPI, VPI, UPISAV,
! VPISAV, the potential-temperature state AADJ, BADJ, PLOGADJ, PL, OM, TL,
! TETA and the layer masses YM; TAUREF (radcommon_h), TAUCUM (standard_h), and
! the per-level aerosol optics of fccsave_h.  Arrays indexed by (J, I) stay
! where they are, and the per-J accumulators are added to in the same column
! order in both passes.  ALS keeps across pass 1 the value it had before it
! (COMP3 sets it with the sun up only).
!
! COMP3's per-step loop for TEMPGR (code/comp3.f:311-351: DSOLFLUX from the
! stored DETAU and the current sun, DNVFLUX = DNDIFFV + DIRECTSOL) goes the same
! way: mgcm_comp3_tempgr_record keeps each column's ACOSZ, and
! mgcm_comp3_tempgr(SOL, DNVFLUX) computes every column's DNVFLUX in one call.
!
! With the environment variable MGCM_COMP3_DUMP naming a file, pass 2 writes
! the records of the per-column glue (comp3_rt.f90) in the same order.
module comp3_split_state
  use grid_h
  use radinc_h
  implicit none
  integer, parameter :: NCMAX = L_J * L_I
  ! COMP3 state between the passes, by (J, I)
  real*8, save :: s_qpi(L_LEVELS, NTRACE, L_J, L_I), s_qpisav(L_LEVELS, NTRACE, L_J, L_I)
  real*8, save :: s_qpig(NTRACE, L_J, L_I), s_tetaorg(L_LEVELS, L_J, L_I), s_dustref(L_J, L_I)
  integer, save :: s_idx(6, L_J, L_I)
  real*8, save :: s_scal(3, L_J, L_I)
  integer, parameter :: NLEV_ARR = 17
  real*8, save :: s_lev(L_LEVELS, NLEV_ARR, L_J, L_I)
  real*8, save :: s_tauref(L_LEVELS+1, L_J, L_I)
  real*8, save :: als_before = 0.0d0
  ! radiation inputs of the recorded columns (c = 1..ncol, in pass order)
  integer, save :: ncol = 0
  integer, save :: colmap(L_J, L_I) = 0
  real*8, save :: r_plev(L_LEVELS, NCMAX), r_tlev(L_LEVELS, NCMAX), r_pmid(L_LEVELS, NCMAX)
  real*8, save :: r_tmid(L_LEVELS, NCMAX), r_qh2o(L_LEVELS, NCMAX)
  real*8, save :: r_tauref(L_LEVELS+1, NCMAX), r_taurefcld(L_LEVELS+1, NCMAX)
  real*8, save :: r_als(NCMAX), r_albi(NCMAX), r_acosz(NCMAX), r_sol(L_NSPECTV)
  real*8, save, target :: r_qxv(L_LEVELS+1, L_NSPECTV, NCMAX), r_qsv(L_LEVELS+1, L_NSPECTV, NCMAX)
  real*8, save, target :: r_gv(L_LEVELS+1, L_NSPECTV, NCMAX), r_qxi(L_LEVELS+1, L_NSPECTI, NCMAX)
  real*8, save, target :: r_qsi(L_LEVELS+1, L_NSPECTI, NCMAX), r_gi(L_LEVELS+1, L_NSPECTI, NCMAX)
  real*8, save, target :: r_cqxv(L_LEVELS+1, L_NSPECTV, NCMAX), r_cqsv(L_LEVELS+1, L_NSPECTV, NCMAX)
  real*8, save, target :: r_cgv(L_LEVELS+1, L_NSPECTV, NCMAX), r_cqxi(L_LEVELS+1, L_NSPECTI, NCMAX)
  real*8, save, target :: r_cqsi(L_LEVELS+1, L_NSPECTI, NCMAX), r_cgi(L_LEVELS+1, L_NSPECTI, NCMAX)
  real*8, save, target :: r_qrd(L_LEVELS+1, NCMAX), r_qrc(L_LEVELS+1, NCMAX)
  ! the batch's results (VIS fields of sun-down columns are not used)
  real*8, save :: o_fmnetv(L_NLAYRAD, NCMAX), o_fluxupv(L_NLAYRAD, NCMAX), o_fluxdnv(L_NLAYRAD, NCMAX)
  real*8, save :: o_fmneti(L_NLAYRAD, NCMAX), o_fluxupi(L_NLAYRAD, NCMAX), o_fluxdni(L_NLAYRAD, NCMAX)
  real*8, save :: o_nfluxtopv(NCMAX), o_nfluxtopi(NCMAX), o_diffvt(NCMAX)
  real*8, save :: o_detau(L_NSPECTV, L_NGAUSS, NCMAX), o_cumtau(L_NSPECTV, L_NGAUSS, NCMAX)
  real*8, save :: o_hr(L_NLAYRAD, NCMAX)
  logical, save :: looked = .false., dump_on = .false.
  integer, save :: dump_unit = 0
  ! COMP3's per-step TEMPGR loop (code/comp3.f:311-351): the columns' ACOSZ, in loop order
  integer, save :: ntg = 0
  integer, save :: tg_j(NCMAX), tg_i(NCMAX)
  real*8, save :: tg_acosz(NCMAX)
contains
  subroutine dump_setup()
    character(len=4096) :: path
    integer :: plen, pst
    looked = .true.
    call get_environment_variable('MGCM_COMP3_DUMP', path, plen, pst)
    dump_on = pst == 0 .and. plen > 0
    if (dump_on) open(newunit=dump_unit, file=path(1:plen), access='stream', form='unformatted', status='replace')
  end subroutine dump_setup

  ! pass 2: the inputs COMP3 hands the radiation must be the ones recorded in pass 1
  subroutine check_same(what, a, b, j, i)
    character(len=*), intent(in) :: what
    real*8, intent(in) :: a(:), b(:)
    integer, intent(in) :: j, i
    if (any(transfer(a, 0_8, size(a)) /= transfer(b, 0_8, size(b)))) then
      write(*, '(a, a, a, i4, a, i4)') 'comp3_split: ', what, ' of column J =', j, ', I =', i
      write(*, '(a)') 'comp3_split: differs between the passes (state not carried by comp3_split_save)'
      stop 1
    end if
  end subroutine check_same
end module comp3_split_state

subroutine comp3_split_pass1_begin()
  use comp3_split_state
  use comp3cmn_h, only: ALS
  implicit none
  als_before = ALS
  ncol = 0
  colmap = 0
end subroutine comp3_split_pass1_begin

subroutine comp3_split_pass1_end()
  use comp3_split_state
  use comp3cmn_h, only: ALS
  implicit none
  ALS = als_before
end subroutine comp3_split_pass1_end

subroutine comp3_split_save(j, i, qpi, qpig, tetaorg, qpisav, dustref, plev, tlev, pmid, tmid, qh2o)
  use comp3_split_state
  use comp3cmn_h
  use radcommon_h, only: TAUREF
  use standard_h, only: TAUCUM
  implicit none
  integer, intent(in) :: j, i
  real*8, intent(in) :: qpi(L_LEVELS, NTRACE), qpig(NTRACE), tetaorg(L_LEVELS), qpisav(L_LEVELS, NTRACE), dustref
  real*8, intent(in) :: plev(L_LEVELS), tlev(L_LEVELS), pmid(L_LEVELS), tmid(L_LEVELS), qh2o(L_LEVELS)
  s_qpi(:, :, j, i) = qpi
  s_qpisav(:, :, j, i) = qpisav
  s_qpig(:, j, i) = qpig
  s_tetaorg(:, j, i) = tetaorg
  s_dustref(j, i) = dustref
  s_idx(:, j, i) = (/ICMN, IM1, IP1, JCMN, JM1, JP1/)
  s_scal(:, j, i) = (/ACOSZ, PSAT, TSAT/)
  s_lev(:, 1, j, i) = TAUICET;  s_lev(:, 2, j, i) = TAUICEL
  s_lev(:, 3, j, i) = UPI;      s_lev(:, 4, j, i) = VPI
  s_lev(:, 5, j, i) = UPISAV;   s_lev(:, 6, j, i) = VPISAV
  s_lev(:, 7, j, i) = AADJ;     s_lev(:, 8, j, i) = BADJ
  s_lev(:, 9, j, i) = PLOGADJ;  s_lev(:, 10, j, i) = PL
  s_lev(:, 11, j, i) = OM;      s_lev(:, 12, j, i) = TL
  s_lev(:, 13, j, i) = TETA;    s_lev(:, 14, j, i) = YM
  s_lev(:, 15, j, i) = TAUCUM
  s_lev(:, 16, j, i) = plev;    s_lev(:, 17, j, i) = tlev
  s_tauref(:, j, i) = TAUREF
  ! pmid, tmid, qh2o are in the radiation record (mgcm_comp3_record, called just before)
  if (colmap(j, i) == 0) stop 'comp3_split: column saved before it was recorded'
  call check_same('PMID', pmid, r_pmid(:, colmap(j, i)), j, i)
end subroutine comp3_split_save

subroutine comp3_split_restore(j, i, qpi, qpig, tetaorg, qpisav, dustref, plev, tlev, pmid, tmid, qh2o)
  use comp3_split_state
  use comp3cmn_h
  use radcommon_h, only: TAUREF
  use standard_h, only: TAUCUM
  use fccsave_h, only: QXVDST, QSVDST, GVDST, QXIDST, QSIDST, GIDST, QXVCLD, QSVCLD, GVCLD, QXICLD, QSICLD, &
                       GICLD, QEXTREFDST, QEXTREFCLD, TAUREFCLD
  implicit none
  integer, intent(in) :: j, i
  real*8, intent(out) :: qpi(L_LEVELS, NTRACE), qpig(NTRACE), tetaorg(L_LEVELS), qpisav(L_LEVELS, NTRACE), dustref
  real*8, intent(out) :: plev(L_LEVELS), tlev(L_LEVELS), pmid(L_LEVELS), tmid(L_LEVELS), qh2o(L_LEVELS)
  integer :: c
  c = colmap(j, i)
  if (c == 0) stop 'comp3_split: column restored that pass 1 did not record'
  qpi = s_qpi(:, :, j, i)
  qpisav = s_qpisav(:, :, j, i)
  qpig = s_qpig(:, j, i)
  tetaorg = s_tetaorg(:, j, i)
  dustref = s_dustref(j, i)
  ICMN = s_idx(1, j, i); IM1 = s_idx(2, j, i); IP1 = s_idx(3, j, i)
  JCMN = s_idx(4, j, i); JM1 = s_idx(5, j, i); JP1 = s_idx(6, j, i)
  ACOSZ = s_scal(1, j, i); PSAT = s_scal(2, j, i); TSAT = s_scal(3, j, i)
  TAUICET = s_lev(:, 1, j, i);  TAUICEL = s_lev(:, 2, j, i)
  UPI = s_lev(:, 3, j, i);      VPI = s_lev(:, 4, j, i)
  UPISAV = s_lev(:, 5, j, i);   VPISAV = s_lev(:, 6, j, i)
  AADJ = s_lev(:, 7, j, i);     BADJ = s_lev(:, 8, j, i)
  PLOGADJ = s_lev(:, 9, j, i);  PL = s_lev(:, 10, j, i)
  OM = s_lev(:, 11, j, i);      TL = s_lev(:, 12, j, i)
  TETA = s_lev(:, 13, j, i);    YM = s_lev(:, 14, j, i)
  TAUCUM = s_lev(:, 15, j, i)
  plev = s_lev(:, 16, j, i);    tlev = s_lev(:, 17, j, i)
  TAUREF = s_tauref(:, j, i)
  pmid = r_pmid(:, c); tmid = r_tmid(:, c); qh2o = r_qh2o(:, c)
  TAUREFCLD = r_taurefcld(:, c)
  QXVDST = r_qxv(:, :, c); QSVDST = r_qsv(:, :, c); GVDST = r_gv(:, :, c)
  QXIDST = r_qxi(:, :, c); QSIDST = r_qsi(:, :, c); GIDST = r_gi(:, :, c)
  QXVCLD = r_cqxv(:, :, c); QSVCLD = r_cqsv(:, :, c); GVCLD = r_cgv(:, :, c)
  QXICLD = r_cqxi(:, :, c); QSICLD = r_cqsi(:, :, c); GICLD = r_cgi(:, :, c)
  QEXTREFDST = r_qrd(:, c); QEXTREFCLD = r_qrc(:, c)
end subroutine comp3_split_restore

! Pass 1: the radiation inputs of column (J, I), with the column's per-level
! aerosol optics as OPT_DST/OPT_CLD (or INI_OPTDST/INI_OPTCLD) left them.
subroutine mgcm_comp3_record(plev, tlev, pmid, tmid, qh2o, tauref, taurefcld, als, albi, acosz, sol, j, i)
  use comp3_split_state
  use fccsave_h, only: QXVDST, QSVDST, GVDST, QXIDST, QSIDST, GIDST, QXVCLD, QSVCLD, GVCLD, QXICLD, QSICLD, &
                       GICLD, QEXTREFDST, QEXTREFCLD
  implicit none
  real*8, intent(in) :: plev(L_LEVELS), tlev(L_LEVELS), pmid(L_LEVELS), tmid(L_LEVELS), qh2o(L_LEVELS)
  real*8, intent(in) :: tauref(L_LEVELS+1), taurefcld(L_LEVELS+1), als, albi, acosz, sol(L_NSPECTV)
  integer, intent(in) :: j, i
  integer :: c
  if (colmap(j, i) /= 0) stop 'comp3_split: column recorded twice in one pass'
  ncol = ncol + 1
  c = ncol
  colmap(j, i) = c
  r_plev(:, c) = plev; r_tlev(:, c) = tlev; r_pmid(:, c) = pmid; r_tmid(:, c) = tmid; r_qh2o(:, c) = qh2o
  r_tauref(:, c) = tauref; r_taurefcld(:, c) = taurefcld
  r_als(c) = als; r_albi(c) = albi; r_acosz(c) = acosz
  if (c == 1) then
    r_sol = sol
  else
    call check_same('SOL', sol, r_sol, j, i)
  end if
  r_qxv(:, :, c) = QXVDST; r_qsv(:, :, c) = QSVDST; r_gv(:, :, c) = GVDST
  r_qxi(:, :, c) = QXIDST; r_qsi(:, :, c) = QSIDST; r_gi(:, :, c) = GIDST
  r_cqxv(:, :, c) = QXVCLD; r_cqsv(:, :, c) = QSVCLD; r_cgv(:, :, c) = GVCLD
  r_cqxi(:, :, c) = QXICLD; r_cqsi(:, :, c) = QSICLD; r_cgi(:, :, c) = GICLD
  r_qrd(:, c) = QEXTREFDST; r_qrc(:, c) = QEXTREFCLD
end subroutine mgcm_comp3_record

! Pass 2, sun up: OPTCV + SFLUXV of column (J, I) from the batch (code/comp3.f:
! 658-675): DETAU(J,I,:,:) as SFLUXV writes it and TAUCUMV(L_LEVELS,:,:) as OPTCV
! leaves it, for COMP3's CUMTAUV loop.
subroutine mgcm_comp3_vis(plev, tlev, pmid, tmid, qh2o, tauref, taurefcld, als, acosz, sol, taucumv, &
                          nfluxtopv, fmnetv, fluxupv, fluxdnv, diffvt, j, i)
  use comp3_split_state
  use radcommon_h, only: DETAU
  implicit none
  real*8, intent(in) :: plev(L_LEVELS), tlev(L_LEVELS), pmid(L_LEVELS), tmid(L_LEVELS), qh2o(L_LEVELS)
  real*8, intent(in) :: tauref(L_LEVELS+1), taurefcld(L_LEVELS+1), als, acosz, sol(L_NSPECTV)
  real*8, intent(inout) :: taucumv(L_LEVELS, L_NSPECTV, L_NGAUSS)
  real*8, intent(out) :: nfluxtopv, fmnetv(L_NLAYRAD), fluxupv(L_NLAYRAD), fluxdnv(L_NLAYRAD), diffvt
  integer, intent(in) :: j, i
  integer :: c
  if (.not. looked) call dump_setup()
  c = colmap(j, i)
  if (c == 0) stop 'comp3_split: VIS call of a column pass 1 did not record'
  call check_same('PLEV', plev, r_plev(:, c), j, i)
  call check_same('TLEV', tlev, r_tlev(:, c), j, i)
  call check_same('PMID/TMID', (/pmid, tmid/), (/r_pmid(:, c), r_tmid(:, c)/), j, i)
  call check_same('QH2O', qh2o, r_qh2o(:, c), j, i)
  call check_same('TAUREF', (/tauref, taurefcld/), (/r_tauref(:, c), r_taurefcld(:, c)/), j, i)
  call check_same('ALS/ACOSZ', (/als, acosz/), (/r_als(c), r_acosz(c)/), j, i)
  call check_same('SOL', sol, r_sol, j, i)
  fmnetv = o_fmnetv(:, c); fluxupv = o_fluxupv(:, c); fluxdnv = o_fluxdnv(:, c)
  nfluxtopv = o_nfluxtopv(c); diffvt = o_diffvt(c)
  taucumv(L_LEVELS, :, :) = o_cumtau(:, :, c)
  DETAU(j, i, :, :) = o_detau(:, :, c)
  if (dump_on) write(dump_unit) int(j, 4), int(i, 4), 1_4, int(3 * L_NLAYRAD + 2 + 2 * L_NSPECTV * L_NGAUSS, 4), &
                                       fmnetv, fluxupv, fluxdnv, nfluxtopv, diffvt, taucumv(L_LEVELS, :, :), &
                                       DETAU(j, i, :, :)
end subroutine mgcm_comp3_vis

! Pass 2: OPTCI + SFLUXI of column (J, I) from the batch (code/comp3.f:724-733).
subroutine mgcm_comp3_ir(plev, tlev, pmid, tmid, qh2o, tauref, taurefcld, albi, nfluxtopi, fmneti, fluxupi, &
                         fluxdni, j, i)
  use comp3_split_state
  implicit none
  real*8, intent(in) :: plev(L_LEVELS), tlev(L_LEVELS), pmid(L_LEVELS), tmid(L_LEVELS), qh2o(L_LEVELS)
  real*8, intent(in) :: tauref(L_LEVELS+1), taurefcld(L_LEVELS+1), albi
  real*8, intent(out) :: nfluxtopi, fmneti(L_NLAYRAD), fluxupi(L_NLAYRAD), fluxdni(L_NLAYRAD)
  integer, intent(in) :: j, i
  integer :: c
  if (.not. looked) call dump_setup()
  c = colmap(j, i)
  if (c == 0) stop 'comp3_split: IR call of a column pass 1 did not record'
  call check_same('PLEV', plev, r_plev(:, c), j, i)
  call check_same('TLEV', tlev, r_tlev(:, c), j, i)
  call check_same('PMID/TMID', (/pmid, tmid/), (/r_pmid(:, c), r_tmid(:, c)/), j, i)
  call check_same('QH2O', qh2o, r_qh2o(:, c), j, i)
  call check_same('TAUREF', (/tauref, taurefcld/), (/r_tauref(:, c), r_taurefcld(:, c)/), j, i)
  call check_same('ALBI', (/albi/), (/r_albi(c)/), j, i)
  fmneti = o_fmneti(:, c); fluxupi = o_fluxupi(:, c); fluxdni = o_fluxdni(:, c)
  nfluxtopi = o_nfluxtopi(c)
  if (dump_on) write(dump_unit) int(j, 4), int(i, 4), 2_4, int(3 * L_NLAYRAD + 1, 4), fmneti, fluxupi, &
                                       fluxdni, nfluxtopi
end subroutine mgcm_comp3_ir

! COMP3's TEMPGR loop, per column: its ACOSZ (code/comp3.f:320-323).
subroutine mgcm_comp3_tempgr_record(acosz, j, i)
  use comp3_split_state
  implicit none
  real*8, intent(in) :: acosz
  integer, intent(in) :: j, i
  if (j == 1 .and. i == 1) ntg = 0
  ntg = ntg + 1
  tg_j(ntg) = j
  tg_i(ntg) = i
  tg_acosz(ntg) = acosz
end subroutine mgcm_comp3_tempgr_record
