C=======================================================================
C  Drop-in replacement of the reference's velocity correction:
C      subroutine uvw_adv(ivar1,diu,rp,dkw,u,u_int,rho,p,wk,offwork,step)
C  (reference code/uvw_adv.F:5, with code/uvw_adv_block.F), compiled in
C  place of code/uvw_adv.F and code/uvw_adv_block.F inside LESter's own
C  build.  u = u_int - 1.5 dtime grad p / rho_face on the component's
C  level-1 range, then the interface/periodic swap of u, on the GPU of
C  the run's liblmgx handle (shared with the drop-in lmg_solve).
C  wk, offwork and step are not used.
C=======================================================================

      subroutine uvw_adv(ivar1,diu,rp,dkw,u,u_int,rho,p,wk,offwork,
     &                   step)

      use iso_c_binding
      use lmgx_mod

      implicit none

#include "dim.inc"
#include "cb.control.inc"

      integer ivar1, step
      real    diu(*), rp(*), dkw(*)
      real    u(*), u_int(*), rho(*), p(*)
      integer offwork(*)
      real    wk(*)

      type(c_ptr) :: h
      integer(c_int) :: nlev, rc

      call lmgx_attach(h, nlev)
      rc = lmgx_uvw_adv(h, ivar1, diu, rp, dkw, u, u_int, rho, p,
     &                  dtime)
      if (rc .ne. 0) then
        write(*,*) 'uvw_adv (liblmgx): ', lmgx_last_error_str()
        stop 'uvw_adv: liblmgx failed'
      end if

      return
      end
