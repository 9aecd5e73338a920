C=======================================================================
C TEST INFRASTRUCTURE ONLY.  Driver that runs the reference LESter
C pressure path (compiled from the reference sources where they lie)
C on a problem described by a binary input file, and dumps everything
C the CPU restatement and the GPU library are checked against.
C
C It mirrors the set-up order of the reference's main.F:165-239 and
C step.F:114-129,222,353-380:
C   set_offset -> set_const -> set_loop_bound(SV_P) -> set_interfaces
C   -> mg_xyz(x,y,z) -> comp_metric -> coeff_p -> sour_p
C   -> lmg_solve (one call per V-cycle, norms logged after each)
C   -> extrap_p -> p_ref_subtract -> uvw_adv (U, V, W)
C   -> (optional) file_3 / read_3 of p
C The commons are filled directly instead of through input.dat /
C block.dat / x.bin, since only their values matter to the path.
C
C Input (stream, little endian; i4 = int32, r8 = float64):
C   i4  magic(=4711), numbl, nummgl_req, meth_cycle, numpre, numpost,
C       numcits, ncycles, coeff_src, src_src, do_post, ntime,
C       meth_smooth (1 = sporlx, 2 = sporlx2; smooth.F:62-73)
C   r8  relax_pmg, relax_pmgc, tolf, dtime
C   per block: i4 nijkp1(3), bc(6), bl2bl(6), bl2side(6)
C   per block: r8 x(2:nip1,2:njp1), y(2:nip1,2:njp1), z(2:nkp1)
C   coeff_src=1: r8 ap,aw,ae,as,an,al,ar  (ntot each, mg layout)
C   src_src=0  : r8 u,v,w,rho,rhonp1 (n1 each)   src_src=1: r8 b (n1)
C   r8  f (n1)   initial guess on level 1
C Output: see the write statements at the end; do_post < 0 writes only
C   i4 4713, ncycles, nummgl; r8 tsolve; r8 hist(4,nummgl,ncycles).
C   After hist the full output holds the reference's global_sum
C   (global_sum.F:5-97) of the final f, b and residu on every level,
C   r8 gsum(3,3,nummgl) = (sum,max,min) x (f,b,residu) x level, and the
C   per-block results of global_sum_block (global_sum.F:105-147),
C   r8 gsb(3,3,numbl,nummgl).
C=======================================================================

      program ref_driver

      implicit none

#include "mpif.h"
#include "dim.inc"
#include "les.inc"
#include "bc_type.inc"
#include "cb.lgrid.inc"
#include "cb.lcontrol.inc"
#include "cb.control.inc"
#include "cb.comm.inc"
#include "cb.gridp.inc"
#include "cb.address.inc"

      integer ierr, magic, nreq, ncycles, coeff_src, src_src, do_post
      integer ntime, itime, msmooth
      integer iblk, is, ilev, n1, ntot, n2d, n1d, nxy, nz, icyc
      integer nip1, njp1, nkp1, i, j, ijk, off, lwork, lgeo
      integer ijkse(6)
      real    t0, t1, tsolve
      character*256 fin, fout, fdump
      character*32 fnb, fnt, fnn

      real, allocatable :: ap(:), aw(:), ae(:), as(:), an(:), al(:),
     &     ar(:), f(:), b(:), res(:), f0(:)
      real, allocatable :: u(:), v(:), w(:), rho(:), rhonp1(:),
     &     drho(:), srhot(:), srhoux(:), gia(:), gja(:),
     &     uo(:), vo(:), wo(:), frb(:), frt(:), frn(:)
      real, allocatable :: x(:), y(:), z(:), work(:), geo(:)
      real, allocatable :: m2(:,:), dkp(:), dkw(:)
      integer, allocatable :: pnti(:,:)
      real, allocatable :: hist(:,:,:)
      real, allocatable :: gsum(:,:,:), gsb(:,:,:,:)
      integer iwork(16)

      call MPI_INIT(ierr)
      call get_command_argument(1, fin)
      call get_command_argument(2, fout)

      open(unit=10, file=fin, access='stream', form='unformatted',
     &     status='old')
      read(10) magic, numbl, nreq, meth_cycle, numpre, numpost,
     &     numcits, ncycles, coeff_src, src_src, do_post, ntime,
     &     msmooth
      if (magic .ne. 4711) stop 'bad magic'
      read(10) relax_pmg, relax_pmgc, tolf, dtime

C COMMUNICATION AND CONTROL COMMONS (ONE RANK OWNS EVERY BLOCK)

      pe = 0
      numpe = 1
      master = .true.
      numbl_total = numbl
      gotrad = .false.
      nummgl = nreq
      meth_smooth = msmooth
      meth_csolve = 0
      meth_rest = 1
      meth_prol = 1
      psolver = 101
      nummgits = 1
      debugging = .false.

      do iblk = 1,numbl
        read(10) (bl_nijkp1(i,iblk),i=1,3), (bl_bc(i,iblk),i=1,6),
     &           (bl2bl(i,iblk),i=1,6), (bl2side(i,iblk),i=1,6)
        bl2pe(iblk) = 0
        do i = 1,3
          bl_nijkp1_master(i,iblk) = bl_nijkp1(i,iblk)
        end do
      end do

      call set_offset()
      call set_const()
      call set_loop_bound(SV_P,1,nummgl,1,numbl,2,ijkse)

C TOTAL SIZES OF THE MULTI-LEVEL ARRAYS

      ntot = mg_off(numbl,nummgl) + mg_nijkp1(1,numbl,nummgl)*
     &     mg_nijkp1(2,numbl,nummgl)*mg_nijkp1(3,numbl,nummgl)
      n1 = mg_off(numbl,1) + mg_nijkp1(1,numbl,1)*
     &     mg_nijkp1(2,numbl,1)*mg_nijkp1(3,numbl,1)
      n2d = mg_off_2d_ij(numbl,nummgl) + mg_nijkp1(1,numbl,nummgl)*
     &     mg_nijkp1(2,numbl,nummgl)
      n1d = mg_off_1d_k(numbl,nummgl) + mg_nijkp1(3,numbl,nummgl)
      nxy = mg_off_2d_xy(numbl,nummgl) +
     &     (mg_nijkp1(1,numbl,nummgl)+1)*(mg_nijkp1(2,numbl,nummgl)+1)
      nz = mg_off_1d_z(numbl,nummgl) + mg_nijkp1(3,numbl,nummgl) + 1

C     ra(jw1a) is both the swap_sc buffer and prol_block's cr scratch
C     (prolongation.F:64-67), so it must hold a whole fine block too
      lwork = maxijk3 + 8*numbl*(maxijk2+8) + 4096
      lrwk = lwork
      lrwk2d = lwork
      liwk = 16
      jw1a = 1
      iwk1a = 1
      allocate(work(lwork))
      work = 0.

      call set_interfaces()

C GEOMETRY: LEVEL-1 VERTICES, THEN THE REFERENCE BUILDS EVERY LEVEL

      allocate(x(nxy), y(nxy), z(nz))
      x = 0.
      y = 0.
      z = 0.
      do iblk = 1,numbl
        nip1 = mg_nijkp1(1,iblk,1)
        njp1 = mg_nijkp1(2,iblk,1)
        nkp1 = mg_nijkp1(3,iblk,1)
        off = mg_off_2d_xy(iblk,1)
        read(10) ((x(off + i + (nip1+1)*(j-1)), i=2,nip1), j=2,njp1)
        read(10) ((y(off + i + (nip1+1)*(j-1)), i=2,nip1), j=2,njp1)
        off = mg_off_1d_z(iblk,1)
        read(10) (z(off + i), i=2,nkp1)
      end do
      call mg_xyz('x',x,work)
      call mg_xyz('y',y,work)
      call mg_xyz('z',z,work)

      allocate(m2(n2d,35), dkp(n1d), dkw(n1d))
      lgeo = 64*(maxijk1_mg+8)*6*numbl + 4096
      allocate(geo(lgeo), pnti(6,numbl))
      geo = 0.
      pnti = 0
      call comp_metric(x,y,z,
     &  m2(1,1),m2(1,2),m2(1,3),m2(1,4),m2(1,5),m2(1,6),m2(1,7),
     &  m2(1,8),m2(1,9),m2(1,10),m2(1,11),m2(1,12),m2(1,13),m2(1,14),
     &  m2(1,15),m2(1,16),m2(1,17),m2(1,18),
     &  m2(1,19),m2(1,20),m2(1,21),m2(1,22),
     &  m2(1,23),m2(1,24),m2(1,25),m2(1,26),
     &  m2(1,27),m2(1,28),m2(1,29),
     &  m2(1,30),m2(1,31),m2(1,32),m2(1,33),m2(1,34),m2(1,35),
     &  dkp,dkw,geo,pnti)

C COEFFICIENTS: THE REFERENCE'S OWN, OR THE ONES IN THE INPUT FILE

      allocate(ap(ntot), aw(ntot), ae(ntot), as(ntot), an(ntot),
     &     al(ntot), ar(ntot), f(ntot), b(ntot), res(ntot), f0(n1))
      ap = 0.
      aw = 0.
      ae = 0.
      as = 0.
      an = 0.
      al = 0.
      ar = 0.
      f = 0.
      b = 0.
      res = 0.
      if (coeff_src .eq. 0) then
C       metric columns: diu=2, djv=10, rdiv=6, rdju=13, didjp=30, rp=27
        call coeff_p(m2(1,2),m2(1,10),m2(1,6),m2(1,13),m2(1,30),
     &       m2(1,27),dkp,dkw,ap,aw,ae,as,an,al,ar)
      else
        read(10) ap, aw, ae, as, an, al, ar
      end if

C SOURCE

      allocate(u(n1), v(n1), w(n1), rho(n1), rhonp1(n1), drho(n1),
     &     srhot(n1), srhoux(n1), gia(maxijk1+2),
     &     gja((maxijk1+2)*(maxijk1+2)))
      u = 0.
      v = 0.
      w = 0.
      rho = 1.
      rhonp1 = 1.
      if (src_src .eq. 0) then
        read(10) u, v, w, rho, rhonp1
C       sour_p writes the level-1 part of b; columns rdidjp=31
        call sour_p(m2(1,2),m2(1,10),m2(1,6),m2(1,13),m2(1,30),dkp,
     &       m2(1,31),dkw,u,v,w,rho,rhonp1,drho,b,srhot,srhoux,
     &       gia,gja)
      else
        read(10) (b(i), i=1,n1)
      end if
      read(10) f0
      close(10)

C SOLVE: ntime REPEATS FROM THE SAME INITIAL GUESS (TIMING); THE LAST
C REPEAT IS THE ONE DUMPED.  jumble PERSISTS ACROSS CALLS, AS IN LESter,
C SO ONLY THE FIRST REPEAT IS KEPT FOR PARITY (ntime=1 FOR GOLDENS).

      allocate(hist(4,nummgl,ncycles))
      hist = 0.
      tsolve = 0.
      do itime = 1,max(ntime,1)
        do i = 1,ntot
          f(i) = 0.
        end do
        do i = 1,n1
          f(i) = f0(i)
        end do
        t0 = MPI_WTIME()
        do icyc = 1,ncycles
          call lmg_solve(ap,aw,ae,as,an,al,ar,f,b,res,iwork,work)
          do ilev = 1,nummgl
            hist(1,ilev,icyc) = rnorm(ilev)
            hist(2,ilev,icyc) = fnorm(ilev)
            hist(3,ilev,icyc) = enorm(ilev)
            hist(4,ilev,icyc) = vnorm(ilev)
          end do
        end do
        t1 = MPI_WTIME()
        tsolve = tsolve + (t1 - t0)
      end do

C TIMING ONLY (do_post < 0): THE SOLVE TIME AND THE NORMS, NO FIELDS

      if (do_post .lt. 0) then
        open(unit=11, file=fout, access='stream', form='unformatted',
     &       status='replace')
        write(11) 4713, ncycles, nummgl
        write(11) tsolve
        write(11) hist
        close(11)
        call MPI_FINALIZE(ierr)
        stop
      end if

      open(unit=11, file=fout, access='stream', form='unformatted',
     &     status='replace')
      write(11) 4712, numbl, nummgl, ntot, n1, n2d, n1d, nxy, nz,
     &     ncycles
      write(11) (((mg_nijkp1(i,iblk,ilev),i=1,3),iblk=1,numbl),
     &     ilev=1,nummgl)
      write(11) ((mg_off(iblk,ilev),iblk=1,numbl),ilev=1,nummgl)
      write(11) ((mg_off_2d_ij(iblk,ilev),iblk=1,numbl),ilev=1,nummgl)
      write(11) ((mg_off_1d_k(iblk,ilev),iblk=1,numbl),ilev=1,nummgl)
      write(11) ((mg_off_2d_xy(iblk,ilev),iblk=1,numbl),ilev=1,nummgl)
      write(11) ((mg_off_1d_z(iblk,ilev),iblk=1,numbl),ilev=1,nummgl)
      write(11) (((nijkse(i,iblk,ilev),i=1,6),iblk=1,numbl),
     &     ilev=1,nummgl)
      write(11) (npts_all(ilev),ilev=1,nummgl)
      write(11) tsolve
      write(11) x, y, z
      write(11) m2, dkp, dkw
      write(11) ap, aw, ae, as, an, al, ar
      write(11) b, f, res
      write(11) hist

C GLOBAL SUMS OF THE FINAL FIELDS, BEFORE THE POST-SOLVE STEPS

      allocate(gsum(3,3,nummgl), gsb(3,3,numbl,nummgl))
      call drv_gsums(f, gsum(1,1,1), gsb(1,1,1,1))
      call drv_gsums(b, gsum(1,2,1), gsb(1,2,1,1))
      call drv_gsums(res, gsum(1,3,1), gsb(1,3,1,1))
      write(11) gsum, gsb

C POST-SOLVE: ZERO-GRADIENT HALOS + SWAP, THEN REFERENCE-POINT SHIFT

      if (do_post .ne. 0) then
        call extrap_p(f,work,iwork)
        call p_ref_subtract(f)
      end if
      write(11) (f(i), i=1,n1)

C VELOCITY CORRECTION WITH THE FINAL p (step.F:418-425), WHEN THE
C VELOCITIES WERE GIVEN: u_int = THE INPUT u, v, w; rho = rhonp1;
C THE CORRECTED FIELDS START AS COPIES OF u_int.  METRICS diu=2,
C djv=10, rp=27 (V PASSES djv IN PLACE OF diu, AS step.F:421 DOES)

      if (src_src .eq. 0 .and. do_post .ne. 0) then
        allocate(uo(n1), vo(n1), wo(n1))
        uo = u
        vo = v
        wo = w
        call uvw_adv(SV_U,m2(1,2),m2(1,27),dkw,uo,u,rhonp1,f,
     &       work,iwork,1)
        call uvw_adv(SV_V,m2(1,10),m2(1,27),dkw,vo,v,rhonp1,f,
     &       work,iwork,1)
        call uvw_adv(SV_W,m2(1,2),m2(1,27),dkw,wo,w,rhonp1,f,
     &       work,iwork,1)
        write(11) n1
        write(11) uo, vo, wo
      else
        write(11) 0
      end if

C FIELD DUMPS (step.F:515, main.F:247), WHEN A THIRD ARGUMENT IS GIVEN:
C file_3 OF THE FINAL p IN BOTH FORMATS INTO THE WORKING DIRECTORY
C (p_bin: filefrmt 0, p_txt: filefrmt 1), THEN read_3 OF EACH, AND OF
C A MISSING FILE (WHICH read_3 ANSWERS WITH ZEROS)

      call get_command_argument(3, fdump)
      if (len_trim(fdump) .gt. 0) then
        allocate(frb(n1), frt(n1), frn(n1))
        frb = -1.
        frt = -1.
        frn = -1.
        ierr = 0
C       file names as LESter passes them: blank-padded character*32
C       (namvar, cb.control.inc:43), which read_3's extr_string scans
        fnb = 'p_bin'
        fnt = 'p_txt'
        fnn = 'p_none'
        call file_3(20,fnb,0,f,1,1,ierr)
        call file_3(21,fnt,1,f,1,1,ierr)
        call read_3(22,fnb,0,frb,work,ierr)
        call read_3(23,fnt,1,frt,work,ierr)
        call read_3(24,fnn,0,frn,work,ierr)
        write(11) n1, ierr
        write(11) frb, frt, frn
      else
        write(11) 0, 0
      end if
      close(11)

      call MPI_FINALIZE(ierr)
      end


C=======================================================================
C global_sum (all blocks) and global_sum_block (each block on its own)
C of one multi-level field on every level.  gs(1:3, ., ilev) gets
C (sum, max, min) with a stride of 9 between levels; gsb(1:3, ., iblk,
C ilev) the block's own results, max and min started afresh per block.
C=======================================================================

      subroutine drv_gsums(fld, gs, gsb)

      implicit none

#include "dim.inc"
#include "cb.lgrid.inc"
#include "cb.lcontrol.inc"
#include "cb.control.inc"

      real    fld(*), gs(9,*), gsb(9,numbl,*)
      integer ilev, iblk, p3
      real    s, tmax, tmin

      do ilev = 1,nummgl
        call global_sum(ilev,1,fld,gs(1,ilev),gs(2,ilev),gs(3,ilev))
        do iblk = 1,numbl
          p3 = mg_off(iblk,ilev) + 1
          tmax = -1.e30
          tmin =  1.e30
          call global_sum_block(fld(p3),mg_nijkp1(1,iblk,ilev),
     &         mg_nijkp1(2,iblk,ilev),mg_nijkp1(3,iblk,ilev),
     &         nijkse(1,iblk,ilev),nijkse(2,iblk,ilev),
     &         nijkse(3,iblk,ilev),nijkse(4,iblk,ilev),
     &         nijkse(5,iblk,ilev),nijkse(6,iblk,ilev),s,tmax,tmin)
          gsb(1,iblk,ilev) = s
          gsb(2,iblk,ilev) = tmax
          gsb(3,iblk,ilev) = tmin
        end do
      end do

      return
      end
