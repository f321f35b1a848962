!  ref_harness.f90 -- bind(C) driver around the reference Fortran radiation
!  routines (TEST INFRASTRUCTURE ONLY).  It is compiled together with the
!  reference's own sources, read where they lie under the reference tree, by
!  oracle/build_ref.sh, into oracle/_ref/libmgcm_ref.so.  It contains no
!  reference code: it sets the module tables, calls OPTCV/SFLUXV/OPTCI/SFLUXI
!  for one column exactly as COMP3 does (code/comp3.f:637-738), and exposes
!  the init routines (SETSPV/SETSPI/SETRAD/LAGINTERP) for table checks.
!
!  The aerosol producers (INITCLD's size grid and optical tables, OPT_DST,
!  OPT_CLD) are exposed the same way: one column at grid point (1,1).
!
!  Per-band Gauss weights: the reference takes one GWEIGHT(17) for all bands.
!  When the caller's weights differ by band, SFLUXV is called once per band
!  with SOL zeroed in the other bands (every VIS term is proportional to
!  F0PI) and SFLUXI once per band with DWNI zeroed in the other bands (every
!  IR accumulation is multiplied by DWNI); the per-band results are summed.

      module ref_harness
      use iso_c_binding
      use grid_h
      use radinc_h
      use radcommon_h, only: planckir, tlimits, tlimiti, ubari
      implicit none
      real*8, save :: detau(L_J,L_I,L_NSPECTV,L_NGAUSS)
      real*8, save, allocatable :: qtrace_h(:,:,:,:)
      contains

!     INITCLD (reads data/*_12bands.dat relative to the working directory)
!     with the dust density INITDT sets (code/initdt.f:38).
      subroutine ref_initcld() bind(C, name="ref_initcld")
      use dtcommon_h, only: dpden_dt
      dpden_dt = 2.5E+3
      call initcld
      end subroutine ref_initcld

!     The size grid and optical tables INITCLD set up (REAL*4 tables promoted).
      subroutine ref_aer_tables(rad, radb, qv_dst, qi_dst, qv_cld, qi_cld,  &
           consts) bind(C, name="ref_aer_tables")
      use cldcommon_h
      use dtcommon_h, only: dpden_dt
      real(c_double), intent(out) :: rad(nbin_rt), radb(nbin_rt+1)
      real(c_double), intent(out) :: qv_dst(nbin_rt, nlonv, 3), qi_dst(nbin_rt, nloni, 3)
      real(c_double), intent(out) :: qv_cld(nratio, nbin_rt, nlonv, 3)
      real(c_double), intent(out) :: qi_cld(nratio, nbin_rt, nloni, 3)
      real(c_double), intent(out) :: consts(6)
      rad = rad_rt
      radb = radb_rt
      qv_dst(:,:,1) = qextv_dst; qv_dst(:,:,2) = qscatv_dst; qv_dst(:,:,3) = gv_dst
      qi_dst(:,:,1) = qexti_dst; qi_dst(:,:,2) = qscati_dst; qi_dst(:,:,3) = gi_dst
      qv_cld(:,:,:,1) = qextv_cld; qv_cld(:,:,:,2) = qscatv_cld; qv_cld(:,:,:,3) = gv_cld
      qi_cld(:,:,:,1) = qexti_cld; qi_cld(:,:,:,2) = qscati_cld; qi_cld(:,:,:,3) = gi_cld
      consts = [athird, dev_dt, dev_ice, dpden_dt, dpden_ice, 0.0d0]
      end subroutine ref_aer_tables

!     DUSTPROFILE over the whole L_J x L_I grid (code/dustprofile.f) with the
!     given namelist values and column opacities TAUTOTJI(L_J, L_I).
      subroutine ref_dust_setup(psf_, ptrop_, rptau_, conrnu_, vary, vpout_, tautot) &
           bind(C, name="ref_dust_setup")
      use standard_h, only: PSF, PTROP, RPTAU, CONRNU, VARY_CONR, VPOUT, TAUTOTJI
      real(c_double), value :: psf_, ptrop_, rptau_, conrnu_, vpout_
      integer(c_int), value :: vary
      real(c_double), intent(in) :: tautot(L_J, L_I)
      PSF = psf_
      PTROP = ptrop_
      RPTAU = rptau_
      CONRNU = conrnu_
      VARY_CONR = vary /= 0
      VPOUT = vpout_
      TAUTOTJI = tautot
      call dustprofile
      end subroutine ref_dust_setup

!     PRDST and TAUDST(j, i, :) of the last ref_dust_setup; then FILLTAUCUM
!     (code/filltaucum.f) for column (j, i) with P(j, i) = p_ji and SIGMA.
      subroutine ref_filltaucum(j, i, p_ji, sig, prdst_out, taudst_out, tauref_out, taucum_out) &
           bind(C, name="ref_filltaucum")
      use standard_h, only: P, SIGMA, TAUCUM, PRDST, TAUDST
      use defines_h, only: L_NPDST
      use radcommon_h, only: TAUREF
      integer(c_int), value :: j, i
      real(c_double), value :: p_ji
      real(c_double), intent(in) :: sig(L_LEVELS)
      real(c_double), intent(out) :: prdst_out(L_NPDST), taudst_out(L_NPDST)
      real(c_double), intent(out) :: tauref_out(L_LEVELS+1), taucum_out(L_LEVELS)
      P(j, i) = p_ji
      SIGMA = sig
      call filltaucum(j, i)
      prdst_out = PRDST
      taudst_out = TAUDST(j, i, :)
      tauref_out = TAUREF
      taucum_out = TAUCUM
      end subroutine ref_filltaucum

!     DSOLFLUX (code/dsolflux.f) of one column's DETAU(7,17).
!     FILLPT (code/fillpt.f90) of one column: the GCM's PL, TL (L_LEVELS), TG,
!     TSTRAT and PTROP -> the radiation grid's PLEV, TLEV, PMID, TMID.
      subroutine ref_fillpt(pl, ptrop_, tg, tstrat, tl, plev, tlev, pmid, tmid) &
           bind(C, name="ref_fillpt")
      real(c_double), intent(in) :: pl(L_LEVELS), tl(L_LEVELS)
      real(c_double), value :: ptrop_, tg, tstrat
      real(c_double), intent(out) :: plev(L_LEVELS), tlev(L_LEVELS), pmid(L_LEVELS), tmid(L_LEVELS)
      real(c_double) :: p_unused
      p_unused = 0.0d0
      call fillpt(pl, p_unused, ptrop_, tg, tstrat, tl, plev, tlev, pmid, tmid)
      end subroutine ref_fillpt

      subroutine ref_dsolflux(sol, ubar0, gweight, fzerov, detau_col, directsol) &
           bind(C, name="ref_dsolflux")
      real(c_double), intent(in) :: sol(L_NSPECTV), gweight(L_NGAUSS), fzerov(L_NSPECTV)
      real(c_double), value :: ubar0
      real(c_double), intent(in) :: detau_col(L_NSPECTV, L_NGAUSS)
      real(c_double), intent(out) :: directsol
      real(c_double), save, allocatable :: d(:,:,:,:)
      real*8 :: sol8(L_NSPECTV), gw8(L_NGAUSS), fz8(L_NSPECTV), u8
      if (.not. allocated(d)) allocate(d(L_J, L_I, L_NSPECTV, L_NGAUSS))
      d = 0.0d0
      d(1, 1, :, :) = detau_col
      sol8 = sol
      gw8 = gweight
      fz8 = fzerov
      u8 = ubar0
      call dsolflux(sol8, u8, gw8, fz8, d, 1, 1, directsol)
      end subroutine ref_dsolflux

!     OPT_DST (cloud = 0) or OPT_CLD (cloud = 1) on one column: tracers
!     QTRACE(L_LAYERS, NTRACE) of the column, PL(L_LEVELS).
      subroutine ref_opt_aer(cloud, qtr, pl, qxv, qxi, qsv, qsi, gv, gi,  &
           qextref, tauref, tau2) bind(C, name="ref_opt_aer")
      use cldcommon_h, only: taudust, taucloud
      integer(c_int), value :: cloud
      real(c_double), intent(in) :: qtr(L_LAYERS, NTRACE), pl(L_LEVELS)
      real(c_double), intent(out) :: qxv(L_LEVELS+1, L_NSPECTV), qsv(L_LEVELS+1, L_NSPECTV)
      real(c_double), intent(out) :: gv(L_LEVELS+1, L_NSPECTV)
      real(c_double), intent(out) :: qxi(L_LEVELS+1, L_NSPECTI), qsi(L_LEVELS+1, L_NSPECTI)
      real(c_double), intent(out) :: gi(L_LEVELS+1, L_NSPECTI)
      real(c_double), intent(out) :: qextref(L_LEVELS+1), tauref(L_LEVELS+1), tau2(2)
      if (.not. allocated(qtrace_h)) allocate(qtrace_h(L_JSIZE, L_ISIZE, L_LAYERS, NTRACE))
      qtrace_h = 0.0d0
      qtrace_h(1, 1, :, :) = qtr
      if (cloud == 0) then
        call opt_dst(1, 1, qtrace_h, pl, qxv, qxi, qsv, qsi, gv, gi, qextref, tauref)
        tau2 = taudust(1, 1, :)
      else
        call opt_cld(1, 1, qtrace_h, pl, qxv, qxi, qsv, qsi, gv, gi, qextref, tauref)
        tau2 = taucloud(1, 1, :)
      end if
      end subroutine ref_opt_aer

      subroutine ref_set_planck(p) bind(C, name="ref_set_planck")
      real(c_double), intent(in) :: p(L_NSPECTI, 8501)
      planckir = p
      end subroutine ref_set_planck

      subroutine ref_nlay(n) bind(C, name="ref_nlay")
      integer(c_int), intent(out) :: n
      n = L_LAYERS
      end subroutine ref_nlay

!     One column.  All arrays in the reference's layouts.
      subroutine ref_column(co2v, co2i, pfgasref, tgasref, wrefh2o,     &
           gwv, gwi, fzerov, fzeroi, tauray, dwni, sol,                  &
           qxvdst, qsvdst, gvdst, qxvcld, qsvcld, gvcld,                 &
           qxidst, qsidst, gidst, qxicld, qsicld, gicld,                 &
           qextrefdst, qextrefcld,                                       &
           plev, tlev, pmid, tmid, qh2o, tauref, taurefcld,              &
           als, albi, acosz,                                             &
           fmnetv, fluxupv, fluxdnv, fmneti, fluxupi, fluxdni,           &
           nfluxtopv, nfluxtopi, diffvt, detau_o, cumtauv_o)             &
           bind(C, name="ref_column")
      real(c_double), intent(in) :: co2v(L_NTREF,L_PINT,L_REFH2O,L_NSPECTV,L_NGAUSS)
      real(c_double), intent(in) :: co2i(L_NTREF,L_PINT,L_REFH2O,L_NSPECTI,L_NGAUSS)
      real(c_double), intent(in) :: pfgasref(L_PINT), tgasref(L_NTREF), wrefh2o(L_REFH2O)
      real(c_double), intent(in) :: gwv(L_NGAUSS,L_NSPECTV), gwi(L_NGAUSS,L_NSPECTI)
      real(c_double), intent(in) :: fzerov(L_NSPECTV), fzeroi(L_NSPECTI)
      real(c_double), intent(in) :: tauray(L_NSPECTV), dwni(L_NSPECTI), sol(L_NSPECTV)
      real(c_double), intent(in) :: qxvdst(L_LEVELS+1,L_NSPECTV), qsvdst(L_LEVELS+1,L_NSPECTV)
      real(c_double), intent(in) :: gvdst(L_LEVELS+1,L_NSPECTV), qxvcld(L_LEVELS+1,L_NSPECTV)
      real(c_double), intent(in) :: qsvcld(L_LEVELS+1,L_NSPECTV), gvcld(L_LEVELS+1,L_NSPECTV)
      real(c_double), intent(in) :: qxidst(L_LEVELS+1,L_NSPECTI), qsidst(L_LEVELS+1,L_NSPECTI)
      real(c_double), intent(in) :: gidst(L_LEVELS+1,L_NSPECTI), qxicld(L_LEVELS+1,L_NSPECTI)
      real(c_double), intent(in) :: qsicld(L_LEVELS+1,L_NSPECTI), gicld(L_LEVELS+1,L_NSPECTI)
      real(c_double), intent(in) :: qextrefdst(L_LEVELS+1), qextrefcld(L_LEVELS+1)
      real(c_double), intent(in) :: plev(L_LEVELS), tlev(L_LEVELS), pmid(L_LEVELS)
      real(c_double), intent(in) :: tmid(L_LEVELS), qh2o(L_LEVELS)
      real(c_double), intent(in) :: tauref(L_LEVELS+1), taurefcld(L_LEVELS+1)
      real(c_double), value :: als, albi, acosz
      real(c_double), intent(out) :: fmnetv(L_NLAYRAD), fluxupv(L_NLAYRAD), fluxdnv(L_NLAYRAD)
      real(c_double), intent(out) :: fmneti(L_NLAYRAD), fluxupi(L_NLAYRAD), fluxdni(L_NLAYRAD)
      real(c_double), intent(out) :: nfluxtopv, nfluxtopi, diffvt
      real(c_double), intent(inout) :: detau_o(L_NSPECTV,L_NGAUSS), cumtauv_o(L_NSPECTV,L_NGAUSS)

      real*8 :: DTAUV(L_NLAYRAD,L_NSPECTV,L_NGAUSS), TAUV(L_NLEVRAD,L_NSPECTV,L_NGAUSS)
      real*8 :: TAUCUMV(L_LEVELS,L_NSPECTV,L_NGAUSS), WBARV(L_NLAYRAD,L_NSPECTV,L_NGAUSS)
      real*8 :: COSBV(L_NLAYRAD,L_NSPECTV,L_NGAUSS), TAUGSURF(L_NSPECTV,L_NGAUSS-1)
      real*8 :: DTAUI(L_NLAYRAD,L_NSPECTI,L_NGAUSS), TAUCUMI(L_LEVELS,L_NSPECTI,L_NGAUSS)
      real*8 :: WBARI(L_NLAYRAD,L_NSPECTI,L_NGAUSS), COSBI(L_NLAYRAD,L_NSPECTI,L_NGAUSS)
      real*8 :: TAUGSURFI(L_NSPECTI,L_NGAUSS-1)
      real*8 :: taurefw(L_LEVELS+1), taurefcldw(L_LEVELS+1)
      real*8 :: qxvdstw(L_LEVELS+1,L_NSPECTV), qxidstw(L_LEVELS+1,L_NSPECTI)
      real*8 :: dwnv(L_NSPECTV), wnoi(L_NSPECTI)
      real*8 :: fmv(L_NLAYRAD), fuv(L_NLAYRAD), fdv(L_NLAYRAD), ntv, dvt
      real*8 :: fmi(L_NLAYRAD), fui(L_NLAYRAD), fdi(L_NLAYRAD), nti
      real*8 :: solb(L_NSPECTV), dwnib(L_NSPECTI), gw(L_NGAUSS)
      logical :: same_v, same_i
      integer :: nw, ng, b

      do ng = 1, L_NGAUSS
        do nw = 1, L_NSPECTV
          detau(1,1,nw,ng) = detau_o(nw,ng)
        end do
      end do
      taurefw = tauref
      taurefcldw = taurefcld
      qxvdstw = qxvdst
      qxidstw = qxidst
      dwnv = 0.0d0
      wnoi = 0.0d0

      same_v = .true.
      same_i = .true.
      do nw = 2, L_NSPECTV
        if (any(gwv(:,nw) .ne. gwv(:,1))) same_v = .false.
      end do
      do nw = 2, L_NSPECTI
        if (any(gwi(:,nw) .ne. gwi(:,1))) same_i = .false.
      end do

      if (acosz .ge. 1.0e-5) then
        call optcv(DTAUV,TAUV,TAUCUMV,CO2V,PLEV,PFGASREF,                &
                   TGASREF,QXVDSTW,QSVDST,GVDST,WBARV,COSBV,             &
                   TAURAY,TAUREFW,TMID,PMID,TAUGSURF,QH2O,               &
                   WREFH2O,                                              &
                   QEXTREFCLD,TAUREFCLDW,QXVCLD,QSVCLD,GVCLD)
        do nw = 1, L_NSPECTV
          do ng = 1, L_NGAUSS
            cumtauv_o(nw,ng) = TAUCUMV(L_LEVELS,nw,ng)
          end do
        end do
        if (same_v) then
          call sfluxv(DTAUV,TAUV,TAUCUMV,ALS,DWNV,WBARV,COSBV,           &
                      ACOSZ,SOL,GWV(:,1),NFLUXTOPV,FMNETV,               &
                      fluxupv,fluxdnv,diffvt,fzerov,taugsurf,            &
                      detau,1,1)
        else
          nfluxtopv = 0.0d0
          fmnetv = 0.0d0
          fluxupv = 0.0d0
          fluxdnv = 0.0d0
          diffvt = 0.0d0
          do b = 1, L_NSPECTV
            solb = 0.0d0
            solb(b) = sol(b)
            gw = gwv(:,b)
            call sfluxv(DTAUV,TAUV,TAUCUMV,ALS,DWNV,WBARV,COSBV,         &
                        ACOSZ,SOLB,GW,NTV,FMV,                           &
                        fuv,fdv,dvt,fzerov,taugsurf,                     &
                        detau,1,1)
            nfluxtopv = nfluxtopv + ntv
            fmnetv = fmnetv + fmv
            fluxupv = fluxupv + fuv
            fluxdnv = fluxdnv + fdv
            diffvt = diffvt + dvt
          end do
        end if
        do ng = 1, L_NGAUSS
          do nw = 1, L_NSPECTV
            detau_o(nw,ng) = detau(1,1,nw,ng)
          end do
        end do
      else
        fmnetv = 0.0d0
        fluxupv = 0.0d0
        fluxdnv = 0.0d0
        nfluxtopv = 0.0d0
        diffvt = 0.0d0
      end if

      call optci(DTAUI,TAUCUMI,CO2I,PLEV,PFGASREF,TGASREF,                &
                 QEXTREFDST,QXIDSTW,QSIDST,GIDST,COSBI,WBARI,             &
                 TAUREFW,TMID,PMID,TAUGSURFI,QH2O,WREFH2O,                &
                 QEXTREFCLD,TAUREFCLDW,QXICLD,QSICLD,GICLD)
      if (same_i) then
        call sfluxi(PLEV,TLEV,DTAUI,TAUCUMI,UBARI,ALBI,WNOI,DWNI,         &
                    COSBI,WBARI,GWI(:,1),NFLUXTOPI,FMNETI,                &
                    FLUXUPI,FLUXDNI,FZEROI,TAUGSURFI)
      else
        nfluxtopi = 0.0d0
        fmneti = 0.0d0
        fluxupi = 0.0d0
        fluxdni = 0.0d0
        do b = 1, L_NSPECTI
          dwnib = 0.0d0
          dwnib(b) = dwni(b)
          gw = gwi(:,b)
          call sfluxi(PLEV,TLEV,DTAUI,TAUCUMI,UBARI,ALBI,WNOI,DWNIB,      &
                      COSBI,WBARI,GW,NTI,FMI,                             &
                      FUI,FDI,FZEROI,TAUGSURFI)
          nfluxtopi = nfluxtopi + nti
          fmneti = fmneti + fmi
          fluxupi = fluxupi + fui
          fluxdni = fluxdni + fdi
        end do
      end if
      end subroutine ref_column

!     A grid of ncol columns in one call (the per-column loop of COMP3 over
!     the grid, code/comp3.f:637-738), so that the reference's own cost can be
!     timed without per-call overhead.  Arrays: one column after another.
      subroutine ref_grid(ncol, co2v, co2i, pfgasref, tgasref, wrefh2o,  &
           gwv, gwi, fzerov, fzeroi, tauray, dwni, sol,                  &
           qxvdst, qsvdst, gvdst, qxvcld, qsvcld, gvcld,                 &
           qxidst, qsidst, gidst, qxicld, qsicld, gicld,                 &
           qextrefdst, qextrefcld,                                       &
           plev, tlev, pmid, tmid, qh2o, tauref, taurefcld,              &
           als, albi, acosz,                                             &
           fmnetv, fluxupv, fluxdnv, fmneti, fluxupi, fluxdni,           &
           nfluxtopv, nfluxtopi, diffvt, detau_o, cumtauv_o)             &
           bind(C, name="ref_grid")
      integer(c_int), value :: ncol
      real(c_double), intent(in) :: co2v(L_NTREF,L_PINT,L_REFH2O,L_NSPECTV,L_NGAUSS)
      real(c_double), intent(in) :: co2i(L_NTREF,L_PINT,L_REFH2O,L_NSPECTI,L_NGAUSS)
      real(c_double), intent(in) :: pfgasref(L_PINT), tgasref(L_NTREF), wrefh2o(L_REFH2O)
      real(c_double), intent(in) :: gwv(L_NGAUSS,L_NSPECTV), gwi(L_NGAUSS,L_NSPECTI)
      real(c_double), intent(in) :: fzerov(L_NSPECTV), fzeroi(L_NSPECTI)
      real(c_double), intent(in) :: tauray(L_NSPECTV), dwni(L_NSPECTI), sol(L_NSPECTV)
      real(c_double), intent(in) :: qxvdst(L_LEVELS+1,L_NSPECTV), qsvdst(L_LEVELS+1,L_NSPECTV)
      real(c_double), intent(in) :: gvdst(L_LEVELS+1,L_NSPECTV), qxvcld(L_LEVELS+1,L_NSPECTV)
      real(c_double), intent(in) :: qsvcld(L_LEVELS+1,L_NSPECTV), gvcld(L_LEVELS+1,L_NSPECTV)
      real(c_double), intent(in) :: qxidst(L_LEVELS+1,L_NSPECTI), qsidst(L_LEVELS+1,L_NSPECTI)
      real(c_double), intent(in) :: gidst(L_LEVELS+1,L_NSPECTI), qxicld(L_LEVELS+1,L_NSPECTI)
      real(c_double), intent(in) :: qsicld(L_LEVELS+1,L_NSPECTI), gicld(L_LEVELS+1,L_NSPECTI)
      real(c_double), intent(in) :: qextrefdst(L_LEVELS+1), qextrefcld(L_LEVELS+1)
      real(c_double), intent(in) :: plev(L_LEVELS,ncol), tlev(L_LEVELS,ncol), pmid(L_LEVELS,ncol)
      real(c_double), intent(in) :: tmid(L_LEVELS,ncol), qh2o(L_LEVELS,ncol)
      real(c_double), intent(in) :: tauref(L_LEVELS+1,ncol), taurefcld(L_LEVELS+1,ncol)
      real(c_double), intent(in) :: als(ncol), albi(ncol), acosz(ncol)
      real(c_double), intent(out) :: fmnetv(L_NLAYRAD,ncol), fluxupv(L_NLAYRAD,ncol)
      real(c_double), intent(out) :: fluxdnv(L_NLAYRAD,ncol), fmneti(L_NLAYRAD,ncol)
      real(c_double), intent(out) :: fluxupi(L_NLAYRAD,ncol), fluxdni(L_NLAYRAD,ncol)
      real(c_double), intent(out) :: nfluxtopv(ncol), nfluxtopi(ncol), diffvt(ncol)
      real(c_double), intent(inout) :: detau_o(L_NSPECTV,L_NGAUSS,ncol)
      real(c_double), intent(inout) :: cumtauv_o(L_NSPECTV,L_NGAUSS,ncol)
      integer :: i
      do i = 1, ncol
        call ref_column(co2v, co2i, pfgasref, tgasref, wrefh2o,         &
             gwv, gwi, fzerov, fzeroi, tauray, dwni, sol,                &
             qxvdst, qsvdst, gvdst, qxvcld, qsvcld, gvcld,               &
             qxidst, qsidst, gidst, qxicld, qsicld, gicld,               &
             qextrefdst, qextrefcld,                                     &
             plev(:,i), tlev(:,i), pmid(:,i), tmid(:,i), qh2o(:,i),      &
             tauref(:,i), taurefcld(:,i),                                &
             als(i), albi(i), acosz(i),                                  &
             fmnetv(:,i), fluxupv(:,i), fluxdnv(:,i),                    &
             fmneti(:,i), fluxupi(:,i), fluxdni(:,i),                    &
             nfluxtopv(i), nfluxtopi(i), diffvt(i),                      &
             detau_o(:,:,i), cumtauv_o(:,:,i))
      end do
      end subroutine ref_grid

!     Init-time tables from the reference's own set-up routines.
      subroutine ref_setspv(wnov, dwnv, wavev, solarf, tauray)          &
           bind(C, name="ref_setspv")
      real(c_double), intent(out) :: wnov(L_NSPECTV), dwnv(L_NSPECTV)
      real(c_double), intent(out) :: wavev(L_NSPECTV), solarf(L_NSPECTV)
      real(c_double), intent(out) :: tauray(L_NSPECTV)
      call setspv(wnov, dwnv, wavev, solarf, tauray)
      end subroutine ref_setspv

      subroutine ref_setspi(wnoi, dwni, wavei) bind(C, name="ref_setspi")
      real(c_double), intent(out) :: wnoi(L_NSPECTI), dwni(L_NSPECTI), wavei(L_NSPECTI)
      call setspi(wnoi, dwni, wavei)
      end subroutine ref_setspi

      subroutine ref_get_planck(p) bind(C, name="ref_get_planck")
      real(c_double), intent(out) :: p(L_NSPECTI, 8501)
      p = planckir
      end subroutine ref_get_planck

!     LAGINTERP reads data/CO2H2O_*_INTEL relative to the working directory.
      subroutine ref_setrad(tgasref, pfgasref, co2v, co2i, qextv, qscatv, &
           wv, gv, qexti, qscati, wi, gi, fzeroi, fzerov)                &
           bind(C, name="ref_setrad")
      real(c_double), intent(out) :: tgasref(L_NTREF), pfgasref(L_PINT)
      real(c_double), intent(out) :: co2v(L_NTREF,L_PINT,L_REFH2O,L_NSPECTV,L_NGAUSS)
      real(c_double), intent(out) :: co2i(L_NTREF,L_PINT,L_REFH2O,L_NSPECTI,L_NGAUSS)
      real(c_double), intent(out) :: qextv(L_NSPECTV), qscatv(L_NSPECTV)
      real(c_double), intent(out) :: wv(L_NSPECTV), gv(L_NSPECTV)
      real(c_double), intent(out) :: qexti(L_NSPECTI), qscati(L_NSPECTI)
      real(c_double), intent(out) :: wi(L_NSPECTI), gi(L_NSPECTI)
      real(c_double), intent(out) :: fzeroi(L_NSPECTI), fzerov(L_NSPECTV)
      call setrad(tgasref, pfgasref, co2v, co2i, qextv, qscatv, wv, gv,   &
                  qexti, qscati, wi, gi, fzeroi, fzerov)
      end subroutine ref_setrad

!     Direct access to the solver pair for per-function checks.
      subroutine ref_dsolver(nl, gama, cp, cm, cpm1, cmm1, e1, e2, e3, e4, &
           btop, bsurf, rsf, xk1, xk2) bind(C, name="ref_dsolver")
      integer(c_int), value :: nl
      real(c_double), intent(in) :: gama(nl), cp(nl), cm(nl), cpm1(nl), cmm1(nl)
      real(c_double), intent(in) :: e1(nl), e2(nl), e3(nl), e4(nl)
      real(c_double), value :: btop, bsurf, rsf
      real(c_double), intent(out) :: xk1(nl), xk2(nl)
      call dsolver(nl, gama, cp, cm, cpm1, cmm1, e1, e2, e3, e4, btop,     &
                   bsurf, rsf, xk1, xk2)
      end subroutine ref_dsolver

      subroutine ref_tpindex(pw, tw, qh2o, pref, tref, wrefh2o, lcoef,   &
           mt, mp, nh2o, wratio) bind(C, name="ref_tpindex")
      real(c_double), value :: pw, tw, qh2o
      real(c_double), intent(in) :: pref(L_PINT), tref(L_NTREF), wrefh2o(L_REFH2O)
      real(c_double), intent(out) :: lcoef(4), wratio
      integer(c_int), intent(out) :: mt, mp, nh2o
      integer :: imt, imp, inh
      real*8 :: pw8, tw8, q8
      pw8 = pw
      tw8 = tw
      q8 = qh2o
      call tpindex(pw8, tw8, q8, pref, tref, wrefh2o, lcoef, imt, imp,    &
                   inh, wratio)
      mt = imt
      mp = imp
      nh2o = inh
      end subroutine ref_tpindex

      end module ref_harness
