C=======================================================================
C  Drop-in replacement of the reference's pressure multigrid:
C      subroutine lmg_solve(ap,aw,ae,as,an,al,ar,f,b,residu,ia,ra)
C  (reference code/lmg_solve.F:5), compiled in place of code/lmg_solve.F
C  inside LESter's own build (it reads LESter's common blocks through
C  LESter's include files).  Topology from /cblgrid1_1/ (numbl,
C  bl_nijkp1, bl_bc, bl2bl, bl2side), multigrid controls from
C  /cbmgdata1/ and /cbmgdata2/ (meth_cycle, nummgl, nummgits, numcits,
C  numpre, numpost, relax_pmg, relax_pmgc, tolf).  On return rnorm,
C  fnorm, enorm, vnorm hold the norms of the last cycle, as the
C  reference leaves them.  ia / ra are not used.
C
C  All blocks must live on this rank (one GPU holds the whole problem).
C  LMGX_SMOOTHER=fast in the environment selects the fast path (zebra
C  line relaxation, FP64); the default is the reference's own schedule
C  (lexicographic SOR with the reference's summation orders), whose
C  results are bitwise those of code/lmg_solve.F.
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
      integer(c_int) :: rc, ncyc
      character(len=16) :: mode
      integer :: l

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
      call get_environment_variable('LMGX_SMOOTHER', mode)
      if (mode(1:4) .eq. 'fast') then
        prm%smoother  = LMGX_SMOOTH_LINE
        prm%precision = LMGX_PREC_FP64
      else
        prm%smoother  = LMGX_SMOOTH_EXACT
      end if

      rc = lmgx_set_operator(h, ap, aw, ae, as, an, al, ar)
      if (rc .eq. 0) then
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
