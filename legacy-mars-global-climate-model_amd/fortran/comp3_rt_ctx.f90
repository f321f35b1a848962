! comp3_rt_ctx.f90 -- the library context of the GCM's radiation call sites on
! the MI355X (the per-column glue comp3_rt.f90 and the two-pass batch
! comp3_rt_batch.f90): set up on first use from the tables RADSETUP left in
! radcommon_h (one GWEIGHT set, repeated per band) and the aerosol constants of
! fccsave_h.  A failing set-up stops the program with the library's message.
module comp3_rt_state
  use iso_c_binding
  implicit none
  type(c_ptr), save :: rtctx = c_null_ptr
end module comp3_rt_state

subroutine comp3_rt_setup()
  use iso_c_binding
  use mgcm_rt_mod
  use comp3_rt_state
  use grid_h
  use radinc_h
  use radcommon_h
  use fccsave_h
  implicit none
  real(c_double) :: gwv(L_NGAUSS, L_NSPECTV), gwi(L_NGAUSS, L_NSPECTI)
  integer(c_int) :: st
  integer :: b
  do b = 1, L_NSPECTV
    gwv(:, b) = GWEIGHT
  end do
  do b = 1, L_NSPECTI
    gwi(:, b) = GWEIGHT
  end do
  call mgcm_rt_setup(CO2V, CO2I, PFGASREF, TGASREF, WREFH2O, gwv, gwi, FZEROV, FZEROI, TAURAY, DWNI, PLANCKIR, &
                     QXVDST(1, :), QSVDST(1, :), GVDST(1, :), QXIDST(1, :), QSIDST(1, :), GIDST(1, :),            &
                     QXVCLD(1, :), QSVCLD(1, :), GVCLD(1, :), QXICLD(1, :), QSICLD(1, :), GICLD(1, :),            &
                     QEXTREFDST(1), QEXTREFCLD(1), rtctx, st)
  if (st /= MGCM_RT_OK) then
    write(*, '(a)') 'comp3_rt: mgcm_rt_setup: '//trim(mgcm_rt_message(st))
    stop 1
  end if
end subroutine comp3_rt_setup

