! comp3_rt.f90 -- COMP3's per-column radiation call sites on the MI355X
! library, for a build of the GCM itself: compiled together with the GCM's
! own modules (grid_h, radinc_h, radcommon_h, fccsave_h) and mgcm_rt_mod, and
! called from COMP3's column loop in place of its four radiation calls.
!
! mgcm_comp3_vis(PLEV, TLEV, PMID, TMID, QH2O, TAUREF, TAUREFCLD, ALS, ACOSZ, SOL,
!                TAUCUMV, NFLUXTOPV, FMNETV, FLUXUPV, FLUXDNV, DIFFVT, J, I)
!   replaces OPTCV + SFLUXV of column (J, I) with the sun up (code/comp3.f:
!   658-675): the VIS part on the device, DETAU(J,I,:,:) written as SFLUXV
!   writes it (code/sfluxv.f90:68-70) and TAUCUMV(L_LEVELS,:,:) as OPTCV leaves
!   it, so COMP3's CUMTAUV loop (code/comp3.f:664-668) runs unchanged.
! mgcm_comp3_ir(PLEV, TLEV, PMID, TMID, QH2O, TAUREF, TAUREFCLD, ALBI,
!               NFLUXTOPI, FMNETI, FLUXUPI, FLUXDNI, J, I)
!   replaces OPTCI + SFLUXI (code/comp3.f:724-733): the IR part on the device.
! Both hand COMP3's per-level aerosol optics of the column -- the fccsave_h
! arrays QXVDST ... GICLD, QEXTREFDST, QEXTREFCLD, which OPT_DST/OPT_CLD fill
! for active dust and clouds and INI_OPTDST/INI_OPTCLD otherwise -- to the
! library through its aerosol argument.  The library context is set up on the
! first call (comp3_rt_ctx.f90).  A failing call stops the program with the
! library's message.
!
! With the environment variable MGCM_COMP3_DUMP naming a file, every call
! appends its outputs to it (stream, native endian): int32 J, I, part (1 VIS,
! 2 IR), n, then n FP64 -- VIS: FMNETV, FLUXUPV, FLUXDNV (L_NLAYRAD),
! NFLUXTOPV, DIFFVT, TAUCUMV(L_LEVELS,:,:) and DETAU(J,I,:,:) (7 x 17 each);
! IR: FMNETI, FLUXUPI, FLUXDNI, NFLUXTOPI.  The drop-in test compares these
! records with the same GCM's run on the reference's own routines.
module comp3_rt_dump
  implicit none
  logical, save :: looked = .false., dump_on = .false.
  integer, save :: dump_unit = 0
contains
  subroutine dump_setup()
    character(len=4096) :: path
    integer :: plen, pst
    looked = .true.
    call get_environment_variable('MGCM_COMP3_DUMP', path, plen, pst)
    dump_on = pst == 0 .and. plen > 0
    if (dump_on) open(newunit=dump_unit, file=path(1:plen), access='stream', form='unformatted', status='replace')
  end subroutine dump_setup
end module comp3_rt_dump

! One column, one part, through mgcm_rt_radiation with the column's per-level aerosol optics.
subroutine comp3_rt_column(part, plev, tlev, pmid, tmid, qh2o, tauref, taurefcld, als, albi, acosz, sol, &
                           fmnet, fluxup, fluxdn, nfluxtop, diffvt, det, cum)
  use iso_c_binding
  use mgcm_rt_mod
  use comp3_rt_state
  use grid_h
  use radinc_h
  use fccsave_h, only: QXVDST, QSVDST, GVDST, QXIDST, QSIDST, GIDST, QXVCLD, QSVCLD, GVCLD, QXICLD, QSICLD, &
                       GICLD, QEXTREFDST, QEXTREFCLD
  implicit none
  integer, intent(in) :: part
  real*8, intent(in) :: plev(L_LEVELS), tlev(L_LEVELS), pmid(L_LEVELS), tmid(L_LEVELS), qh2o(L_LEVELS)
  real*8, intent(in) :: tauref(L_LEVELS+1), taurefcld(L_LEVELS+1), als, albi, acosz, sol(L_NSPECTV)
  real*8, intent(out) :: fmnet(L_NLAYRAD), fluxup(L_NLAYRAD), fluxdn(L_NLAYRAD), nfluxtop, diffvt
  real*8, intent(inout) :: det(L_NSPECTV, L_NGAUSS), cum(L_NSPECTV, L_NGAUSS)
  real(c_double), target :: qxv(L_LEVELS+1, L_NSPECTV), qsv(L_LEVELS+1, L_NSPECTV), gv(L_LEVELS+1, L_NSPECTV)
  real(c_double), target :: qxi(L_LEVELS+1, L_NSPECTI), qsi(L_LEVELS+1, L_NSPECTI), gi(L_LEVELS+1, L_NSPECTI)
  real(c_double), target :: cqxv(L_LEVELS+1, L_NSPECTV), cqsv(L_LEVELS+1, L_NSPECTV), cgv(L_LEVELS+1, L_NSPECTV)
  real(c_double), target :: cqxi(L_LEVELS+1, L_NSPECTI), cqsi(L_LEVELS+1, L_NSPECTI), cgi(L_LEVELS+1, L_NSPECTI)
  real(c_double), target :: qrd(L_LEVELS+1), qrc(L_LEVELS+1)
  type(mgcm_rt_aerosol), target :: aer
  real(c_double) :: pl(L_LEVELS, 1), tl(L_LEVELS, 1), pm(L_LEVELS, 1), tm(L_LEVELS, 1), qh(L_LEVELS, 1)
  real(c_double) :: tr(L_LEVELS+1, 1), tc(L_LEVELS+1, 1), a1(1), a2(1), a3(1)
  real(c_double) :: fv(L_NLAYRAD, 1), uv(L_NLAYRAD, 1), dv(L_NLAYRAD, 1), fi(L_NLAYRAD, 1), ui(L_NLAYRAD, 1)
  real(c_double) :: di(L_NLAYRAD, 1), ntv(1), nti(1), dif(1), hr(L_NLAYRAD, 1)
  real(c_double) :: d3(L_NSPECTV, L_NGAUSS, 1), c3(L_NSPECTV, L_NGAUSS, 1)
  integer(c_int) :: st, mode
  if (.not. c_associated(rtctx)) call comp3_rt_setup()
  qxv = QXVDST; qsv = QSVDST; gv = GVDST; qxi = QXIDST; qsi = QSIDST; gi = GIDST
  cqxv = QXVCLD; cqsv = QSVCLD; cgv = GVCLD; cqxi = QXICLD; cqsi = QSICLD; cgi = GICLD
  qrd = QEXTREFDST
  qrc = QEXTREFCLD
  aer%qxv_dst = c_loc(qxv); aer%qsv_dst = c_loc(qsv); aer%gv_dst = c_loc(gv)
  aer%qxi_dst = c_loc(qxi); aer%qsi_dst = c_loc(qsi); aer%gi_dst = c_loc(gi)
  aer%qxv_cld = c_loc(cqxv); aer%qsv_cld = c_loc(cqsv); aer%gv_cld = c_loc(cgv)
  aer%qxi_cld = c_loc(cqxi); aer%qsi_cld = c_loc(cqsi); aer%gi_cld = c_loc(cgi)
  aer%qextref_dst = c_loc(qrd)
  aer%qextref_cld = c_loc(qrc)
  pl(:, 1) = plev; tl(:, 1) = tlev; pm(:, 1) = pmid; tm(:, 1) = tmid; qh(:, 1) = qh2o
  tr(:, 1) = tauref; tc(:, 1) = taurefcld
  a1(1) = als; a2(1) = albi; a3(1) = acosz
  d3(:, :, 1) = det
  c3(:, :, 1) = cum
  if (part == 1) then
    mode = MGCM_RT_FP64 + MGCM_RT_SKIP_IR
  else
    mode = MGCM_RT_FP64 + MGCM_RT_SKIP_VIS
  end if
  call mgcm_rt_radiation(rtctx, 1, L_LAYERS, pl, tl, pm, tm, qh, tr, tc, a1, a2, a3, sol, fv, uv, dv, fi, ui, di, &
                         ntv, nti, dif, d3, c3, hr, mode, st, aerosol=aer)
  if (st /= MGCM_RT_OK) then
    write(*, '(a)') 'comp3_rt: mgcm_rt_radiation: '//trim(mgcm_rt_message(st))
    stop 1
  end if
  if (part == 1) then
    fmnet = fv(:, 1); fluxup = uv(:, 1); fluxdn = dv(:, 1); nfluxtop = ntv(1); diffvt = dif(1)
    det = d3(:, :, 1)
    cum = c3(:, :, 1)
  else
    fmnet = fi(:, 1); fluxup = ui(:, 1); fluxdn = di(:, 1); nfluxtop = nti(1); diffvt = 0.0d0
  end if
end subroutine comp3_rt_column

subroutine mgcm_comp3_vis(plev, tlev, pmid, tmid, qh2o, tauref, taurefcld, als, acosz, sol, taucumv, &
                          nfluxtopv, fmnetv, fluxupv, fluxdnv, diffvt, j, i)
  use comp3_rt_dump
  use grid_h
  use radinc_h
  use radcommon_h, only: DETAU
  implicit none
  real*8, intent(in) :: plev(L_LEVELS), tlev(L_LEVELS), pmid(L_LEVELS), tmid(L_LEVELS), qh2o(L_LEVELS)
  real*8, intent(in) :: tauref(L_LEVELS+1), taurefcld(L_LEVELS+1), als, acosz, sol(L_NSPECTV)
  real*8, intent(inout) :: taucumv(L_LEVELS, L_NSPECTV, L_NGAUSS)
  real*8, intent(out) :: nfluxtopv, fmnetv(L_NLAYRAD), fluxupv(L_NLAYRAD), fluxdnv(L_NLAYRAD), diffvt
  integer, intent(in) :: j, i
  real*8 :: det(L_NSPECTV, L_NGAUSS), cum(L_NSPECTV, L_NGAUSS)
  if (.not. looked) call dump_setup()
  det = DETAU(j, i, :, :)
  cum = taucumv(L_LEVELS, :, :)
  call comp3_rt_column(1, plev, tlev, pmid, tmid, qh2o, tauref, taurefcld, als, 0.0d0, acosz, sol, &
                       fmnetv, fluxupv, fluxdnv, nfluxtopv, diffvt, det, cum)
  DETAU(j, i, :, :) = det
  taucumv(L_LEVELS, :, :) = cum
  if (dump_on) write(dump_unit) int(j, 4), int(i, 4), 1_4, int(3 * L_NLAYRAD + 2 + 2 * L_NSPECTV * L_NGAUSS, 4), &
                                       fmnetv, fluxupv, fluxdnv, nfluxtopv, diffvt, cum, det
end subroutine mgcm_comp3_vis

subroutine mgcm_comp3_ir(plev, tlev, pmid, tmid, qh2o, tauref, taurefcld, albi, nfluxtopi, fmneti, fluxupi, &
                         fluxdni, j, i)
  use comp3_rt_dump
  use grid_h
  use radinc_h
  implicit none
  real*8, intent(in) :: plev(L_LEVELS), tlev(L_LEVELS), pmid(L_LEVELS), tmid(L_LEVELS), qh2o(L_LEVELS)
  real*8, intent(in) :: tauref(L_LEVELS+1), taurefcld(L_LEVELS+1), albi
  real*8, intent(out) :: nfluxtopi, fmneti(L_NLAYRAD), fluxupi(L_NLAYRAD), fluxdni(L_NLAYRAD)
  integer, intent(in) :: j, i
  real*8 :: det(L_NSPECTV, L_NGAUSS), cum(L_NSPECTV, L_NGAUSS), sol(L_NSPECTV), dif
  if (.not. looked) call dump_setup()
  det = 0.0d0
  cum = 0.0d0
  sol = 0.0d0
  call comp3_rt_column(2, plev, tlev, pmid, tmid, qh2o, tauref, taurefcld, 0.0d0, albi, 0.0d0, sol, &
                       fmneti, fluxupi, fluxdni, nfluxtopi, dif, det, cum)
  if (dump_on) write(dump_unit) int(j, 4), int(i, 4), 2_4, int(3 * L_NLAYRAD + 1, 4), fmneti, fluxupi, &
                                       fluxdni, nfluxtopi
end subroutine mgcm_comp3_ir
