C=======================================================================
C  Drop-in replacement of the reference's pressure multigrid:
C      subroutine lmg_solve(ap,aw,ae,as,an,al,ar,f,b,residu,ia,ra)
C  (reference code/lmg_solve.F:5), compiled in place of code/lmg_solve.F
C  inside LESter's own build (it reads LESter's common blocks through
C  LESter's include files).  Topology from /cblgrid1_1/ (numbl,
C  bl_nijkp1, bl_bc, bl2bl, bl2side), multigrid controls from
C  /cbmgdata1/ and /cbmgdata2/ (meth_cycle, meth_smooth, meth_csolve,
C  nummgl, nummgits, numcits, numpre, numpost, relax_pmg, relax_pmgc,
C  tolf).  On return rnorm, fnorm, enorm, vnorm hold the norms of the
C  last cycle, as the reference leaves them.  ia / ra are not used.
C
C  All blocks must live on this rank (one GPU holds the whole problem).
C
C  Schedule, chosen from the environment at each call:
C    LMGX_SMOOTHER unset or 'exact': the reference's own schedule with
C      its smoother meth_smooth = 1 (sporlx) or 2 (sporlx2); f, b and
C      residu on every level are bitwise those of code/lmg_solve.F.
C    LMGX_SMOOTHER = 'fast': zebra line relaxation with the line
C      smoother's own tuned knobs (over-relaxation 1.35, or LMGX_OMEGA;
C      3 line sweeps per smoothing iteration; coarse levels from level 1;
C      exact coarsest solve); numpre, numpost, nummgits and tolf from the
C      commons.  LMGX_PRECISION = 'mixed' runs it as an FP32 V-cycle
C      inside FP64 iterative refinement (FP64 residual, correction,
C      coarsest solve and norms).  Only level 1 of f and residu is
C      written (the coarse levels live on the GPU in their own layout).
C  A method the library does not have stops the run with a message, as
C  the reference does for meth_smooth (smooth.F:77-79), meth_csolve
C  (solversc.F:41-43) and psolver (solvers.F:35-36).
C
C  The coefficients stay on the GPU between calls (LESter computes them
C  once, code/step.F:359-362): they are uploaded on the first call and
C  again when their arrays or a sampled fingerprint of them change, or
C  after a call of lmgx_coeff_changed (for a caller that edits a few
C  coefficients in place).  LMGX_VERBOSE=1 prints the library's counters
C  (uploads, operator preparations, graph captures, solves) each call.
C=======================================================================

      subroutine lmg_solve(ap,aw,ae,as,an,al,ar,f,b,residu,ia,ra)

      use iso_c_binding
      use lmgx_mod

      implicit none

#include "dim.inc"
#include "cb.lgrid.inc"
#include "cb.control.inc"

      integer   ia(*)
      real      ra(*)
      real      ap(*), aw(*), ae(*), as(*), an(*), al(*), ar(*)
      real      f(*), b(*), residu(*)

      type(c_ptr) :: h
      integer(c_int) :: nlev
      type(lmgx_params) :: prm
      real(c_double), allocatable :: hist(:,:,:)
      integer(c_int) :: rc, ncyc, uploaded
      integer(c_long) :: cnt(4)
      character(len=16) :: mode, prec, verb
      character(len=32) :: om
      integer :: l, ios

      if (psolver .ne. 101) then
        write(*,'(''*** Error: pressure solver '',i4,
     &            '' is not available (lmg_solve on liblmgx)'')')
     &            psolver
        stop
      end if
      if (meth_csolve .ne. 0) then
        write(*,'(''*** Error: coarsest-grid method '',i4,
     &            '' is not available (lmg_solve on liblmgx)'')')
     &            meth_csolve
        stop
      end if

      if (meth_smooth .ne. 1 .and. meth_smooth .ne. 2) then
        write(*,'(''*** Error: smoothing method '',i4,
     &            '' is not available (lmg_solve on liblmgx)'')')
     &            meth_smooth
        stop
      end if

      call lmgx_attach(h, nlev)

      call lmgx_default_params(prm)
      prm%cycle      = meth_cycle
      prm%numpre     = numpre
      prm%numpost    = numpost
      prm%numcits    = numcits
      prm%max_cycles = nummgits
      prm%omega      = relax_pmg
      prm%omega_c    = relax_pmgc
      prm%tolf       = tolf
      prm%rtol       = 0.d0
      mode = ' '
      prec = ' '
      verb = ' '
      call get_environment_variable('LMGX_SMOOTHER', mode)
      call get_environment_variable('LMGX_PRECISION', prec)
      call get_environment_variable('LMGX_VERBOSE', verb)
      if (mode(1:4) .eq. 'fast') then
        prm%smoother        = LMGX_SMOOTH_LINE
        prm%precision       = LMGX_PREC_FP64
        if (prec(1:5) .eq. 'mixed') prm%precision = LMGX_PREC_MIXED
        prm%omega           = 1.35d0
        om = ' '
        call get_environment_variable('LMGX_OMEGA', om)
        if (len_trim(om) .gt. 0) then
          read(om, *, iostat=ios) prm%omega
          if (ios .ne. 0) stop 'lmg_solve: LMGX_OMEGA is not a number'
        end if
        prm%omega_c         = 1.d0
        prm%sweeps_per_iter = 3
        prm%coarse_op       = 1
        prm%coarse_solve    = 1
      else if (meth_smooth .eq. 2) then
        prm%smoother  = LMGX_SMOOTH_SPORLX2
      else
        prm%smoother  = LMGX_SMOOTH_EXACT
      end if

      rc = lmgx_set_operator_cached(h, ap, aw, ae, as, an, al, ar,
     &                              lmgx_force_upload, uploaded)
      if (rc .eq. 0) then
        lmgx_force_upload = 0
        allocate(hist(4, nlev, max(1, nummgits)))
        rc = lmgx_solve(h, f, b, residu, prm, hist, ncyc)
        if (rc .eq. 0 .and. ncyc .gt. 0) then
          do l = 1, nlev
            rnorm(l) = hist(1, l, ncyc)
            fnorm(l) = hist(2, l, ncyc)
            enorm(l) = hist(3, l, ncyc)
            vnorm(l) = hist(4, l, ncyc)
          end do
        end if
        deallocate(hist)
      end if
      if (rc .ne. 0) then
        write(*,*) 'lmg_solve (liblmgx): ', lmgx_last_error_str()
        stop 'lmg_solve: liblmgx failed'
      end if
      if (verb(1:1) .eq. '1') then
        rc = lmgx_counters(h, cnt)
        write(*,'(a,4i10)') ' lmgx counters (uploads, preparations,'//
     &    ' captures, solves):', cnt
      end if

      return
      end


C=======================================================================
C  Marks the coefficients as changed: the next lmg_solve uploads them
C  (and the fast path re-derives its operator) even if the arrays and
C  their sampled fingerprint are those of the last upload.
C=======================================================================

      subroutine lmgx_coeff_changed()

      use lmgx_mod

      implicit none

      lmgx_force_upload = 1

      return
      end


C=======================================================================
C  The run's liblmgx handle (module variable lmgx_shared), created on
C  the first call from LESter's topology commons; nlev = the level count
C  the library settled on.  Used by the drop-ins lmg_solve and uvw_adv.
C=======================================================================

      subroutine lmgx_attach(h, nlev)

      use iso_c_binding
      use lmgx_mod

      implicit none

#include "dim.inc"
#include "cb.lgrid.inc"
#include "cb.control.inc"

      type(c_ptr), intent(out) :: h
      integer(c_int), intent(out) :: nlev

      integer(c_int), save :: nlev_s = 0
      integer(c_int), allocatable, target, save :: tdims(:,:)
      integer(c_int), allocatable, target, save :: tbc(:,:)
      integer(c_int), allocatable, target, save :: tbl(:,:)
      integer(c_int), allocatable, target, save :: tsd(:,:)
      type(lmgx_topology) :: topo
      integer(c_int) :: rc
      integer(c_long) :: ntot, n1
      integer(c_int), allocatable :: dims(:)
      integer(c_long), allocatable :: off(:)

      if (.not. c_associated(lmgx_shared)) then
        allocate(tdims(3,numbl), tbc(6,numbl), tbl(6,numbl),
     &           tsd(6,numbl))
        tdims = bl_nijkp1(:,1:numbl)
        tbc   = bl_bc(:,1:numbl)
        tbl   = bl2bl(:,1:numbl)
        tsd   = bl2side(:,1:numbl)
        topo%nblk    = numbl
        topo%nlev    = nummgl
        topo%nijkp1  = c_loc(tdims)
        topo%bc      = c_loc(tbc)
        topo%bl2bl   = c_loc(tbl)
        topo%bl2side = c_loc(tsd)
        rc = lmgx_init(topo, lmgx_shared)
        if (rc .ne. 0) then
          write(*,*) 'liblmgx: ', lmgx_last_error_str()
          stop 'liblmgx init failed'
        end if
        allocate(dims(3*maxbl*8), off(maxbl*8))
        rc = lmgx_levels(lmgx_shared, nlev_s, ntot, n1, dims, off)
        deallocate(dims, off)
      end if
      h = lmgx_shared
      nlev = nlev_s

      return
      end
