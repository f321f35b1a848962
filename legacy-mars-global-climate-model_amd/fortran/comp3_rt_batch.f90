! comp3_rt_batch.f90 -- comp3_split_radiate on the MI355X: every column that
! COMP3's first pass recorded (comp3_split.f90), VIS and IR, in one
! mgcm_rt_radiation call with the columns' per-level aerosol optics.  Sun-down
! columns come back without VIS results, which COMP3's sun-down branch never
! asks for; DETAU and CUMTAUV go in as the columns hold them.  A failing call
! stops the program with the library's message.
subroutine comp3_split_radiate()
  use iso_c_binding
  use mgcm_rt_mod
  use comp3_rt_state
  use comp3_split_state
  use radcommon_h, only: DETAU
  use fccsave_h, only: CUMTAUV
  implicit none
  type(mgcm_rt_aerosol), target :: aer
  integer(c_int) :: st
  integer :: c, j, i
  if (ncol == 0) return
  if (.not. c_associated(rtctx)) call comp3_rt_setup()
  do i = 1, L_I
    do j = 1, L_J
      c = colmap(j, i)
      if (c /= 0) then
        o_detau(:, :, c) = DETAU(j, i, :, :)
        o_cumtau(:, :, c) = CUMTAUV(j, i, :, :)
      end if
    end do
  end do
  aer%qxv_dst = c_loc(r_qxv); aer%qsv_dst = c_loc(r_qsv); aer%gv_dst = c_loc(r_gv)
  aer%qxi_dst = c_loc(r_qxi); aer%qsi_dst = c_loc(r_qsi); aer%gi_dst = c_loc(r_gi)
  aer%qxv_cld = c_loc(r_cqxv); aer%qsv_cld = c_loc(r_cqsv); aer%gv_cld = c_loc(r_cgv)
  aer%qxi_cld = c_loc(r_cqxi); aer%qsi_cld = c_loc(r_cqsi); aer%gi_cld = c_loc(r_cgi)
  aer%qextref_dst = c_loc(r_qrd)
  aer%qextref_cld = c_loc(r_qrc)
  call mgcm_rt_radiation(rtctx, ncol, L_LAYERS, r_plev, r_tlev, r_pmid, r_tmid, r_qh2o, r_tauref, r_taurefcld, &
                         r_als, r_albi, r_acosz, r_sol, o_fmnetv, o_fluxupv, o_fluxdnv, o_fmneti, o_fluxupi, &
                         o_fluxdni, o_nfluxtopv, o_nfluxtopi, o_diffvt, o_detau, o_cumtau, o_hr, MGCM_RT_FP64, st, &
                         aerosol=aer)
  if (st /= MGCM_RT_OK) then
    write(*, '(a)') 'comp3_rt_batch: mgcm_rt_radiation: '//trim(mgcm_rt_message(st))
    stop 1
  end if
end subroutine comp3_split_radiate

! COMP3's TEMPGR loop (code/comp3.f:311-351) on the MI355X: DSOLFLUX of every
! recorded column from the stored DETAU and its current ACOSZ, and DNVFLUX(J,I) =
! DNDIFFV(J,I) + DIRECTSOL, in one mgcm_rt_dsolflux_host call.
subroutine mgcm_comp3_tempgr(sol, dnvflux)
  use iso_c_binding
  use mgcm_rt_mod
  use comp3_rt_state
  use comp3_split_state
  use radinc_h
  use radcommon_h, only: DETAU
  use standard_h, only: DNDIFFV
  implicit none
  real*8, intent(in) :: sol(L_NSPECTV)
  real*8, intent(inout) :: dnvflux(L_J, L_I)  ! COMP3's DNVFLUX
  real(c_double), allocatable :: det(:, :, :), dif(:), dir(:), dnv(:)
  integer(c_int) :: st
  integer :: c
  if (ntg == 0) return
  if (.not. c_associated(rtctx)) call comp3_rt_setup()
  allocate(det(L_NSPECTV, L_NGAUSS, ntg), dif(ntg), dir(ntg), dnv(ntg))
  do c = 1, ntg
    det(:, :, c) = DETAU(tg_j(c), tg_i(c), :, :)
    dif(c) = DNDIFFV(tg_j(c), tg_i(c))
  end do
  st = mgcm_rt_dsolflux_host(rtctx, int(ntg, c_int), sol, tg_acosz, det, dif, dir, dnv)
  if (st /= MGCM_RT_OK) then
    write(*, '(a)') 'comp3_rt_batch: mgcm_rt_dsolflux_host: '//trim(mgcm_rt_message(st))
    stop 1
  end if
  do c = 1, ntg
    dnvflux(tg_j(c), tg_i(c)) = dnv(c)
  end do
end subroutine mgcm_comp3_tempgr
