!------------------------------------------------------------------------------
! C-callable entry points into the reference's own ModuleFunctions, compiled
! from Software/MOHIDBase1/ModuleFunctions.F90 where it lies (see oracle/Makefile).
!
! TEST INFRASTRUCTURE ONLY: this file is project code that binds the reference
! build so the tests can compare the C oracle with the reference routines
! bit for bit.  It is never linked into the product library.
!
! Built with the reference's production precision (-fdefault-real-8
! -fdefault-double-8, Solutions/mohid-in-linux/compile_mohid.sh:185), so every
! `real` below is an 8-byte real like the reference's.
!------------------------------------------------------------------------------
module ref_mf_bind
    use iso_c_binding
    use ModuleFunctions, only : THOMASZ_NewType2, THOMAS_3D, ComputeAdvection1D_V2, &
                                ComputeAdvection1D_TVD_SuperBee_2,              &
                                ComputeAdvectionFace, OrlanskiCelerity2D,       &
                                T_THOMAS, T_D_E_F, T_VECGW
    implicit none

contains

    !> THOMASZ_NewType2 (ModuleFunctions.F90:4026-4123) on caller arrays that
    !> span Size (sz = ILB,IUB,JLB,JUB,KLB,KUB), column-major.  vecw/vecg
    !> carry the per-thread VEC%W/G scratch in and out (index lo..hi).
    subroutine ref_thomasz(sz, ws, lo, hi, D, E, F, TI, WP, RES, vecw, vecg)      &
               bind(c, name='ref_thomasz')
        integer(c_int), intent(in)        :: sz(6), ws(6)
        integer(c_int), value             :: lo, hi
        type(c_ptr), value                :: D, E, F, TI, WP, RES, vecw, vecg
        real,    pointer :: fD(:,:,:), fE(:,:,:), fF(:,:,:), fTI(:,:,:), fR(:,:,:)
        real,    pointer :: bD(:,:,:), bE(:,:,:), bF(:,:,:), bTI(:,:,:), bR(:,:,:)
        integer, pointer :: fWP(:,:,:), bWP(:,:,:)
        real,    pointer :: fw(:), fg(:)
        type(T_THOMAS), pointer :: th
        integer :: shp(3)

        shp = [sz(2)-sz(1)+1, sz(4)-sz(3)+1, sz(6)-sz(5)+1]
        call c_f_pointer(D,   fD,  shp); bD (sz(1):, sz(3):, sz(5):) => fD
        call c_f_pointer(E,   fE,  shp); bE (sz(1):, sz(3):, sz(5):) => fE
        call c_f_pointer(F,   fF,  shp); bF (sz(1):, sz(3):, sz(5):) => fF
        call c_f_pointer(TI,  fTI, shp); bTI(sz(1):, sz(3):, sz(5):) => fTI
        call c_f_pointer(RES, fR,  shp); bR (sz(1):, sz(3):, sz(5):) => fR
        call c_f_pointer(WP,  fWP, shp); bWP(sz(1):, sz(3):, sz(5):) => fWP
        call c_f_pointer(vecw, fw, [hi-lo+1])
        call c_f_pointer(vecg, fg, [hi-lo+1])

        allocate(th)
        allocate(th%COEF3)
        th%COEF3%D => bD
        th%COEF3%E => bE
        th%COEF3%F => bF
        th%TI      => bTI
        allocate(th%VEC(1))
        allocate(th%VEC(1)%W(lo:hi), th%VEC(1)%G(lo:hi))
        th%VEC(1)%W(lo:hi) = fw
        th%VEC(1)%G(lo:hi) = fg

        call THOMASZ_NewType2(ws(1), ws(2), ws(3), ws(4), ws(5), ws(6), th, bR, bWP)

        fw = th%VEC(1)%W(lo:hi)
        fg = th%VEC(1)%G(lo:hi)
        deallocate(th%VEC(1)%W, th%VEC(1)%G)
        deallocate(th%VEC)
        deallocate(th%COEF3)
        deallocate(th)
    end subroutine ref_thomasz

    !> THOMAS_3D (the T_THOMAS form, THOMAS_3D_NewType, ModuleFunctions.F90:
    !> 3667-3870) on caller arrays that span Size, column-major.  lim =
    !> IJmin, IJmax, JImin, JImax, Kmin, Kmax; vecw/vecg as for ref_thomasz.
    subroutine ref_thomas3d(sz, lim, di, dj, lo, hi, D, E, F, TI, RES, vecw, vecg) &
               bind(c, name='ref_thomas3d')
        integer(c_int), intent(in)        :: sz(6), lim(6)
        integer(c_int), value             :: di, dj, lo, hi
        type(c_ptr), value                :: D, E, F, TI, RES, vecw, vecg
        real,    pointer :: fD(:,:,:), fE(:,:,:), fF(:,:,:), fTI(:,:,:), fR(:,:,:)
        real,    pointer :: bD(:,:,:), bE(:,:,:), bF(:,:,:), bTI(:,:,:), bR(:,:,:)
        real,    pointer :: fw(:), fg(:)
        type(T_THOMAS), pointer :: th
        integer :: shp(3)

        shp = [sz(2)-sz(1)+1, sz(4)-sz(3)+1, sz(6)-sz(5)+1]
        call c_f_pointer(D,   fD,  shp); bD (sz(1):, sz(3):, sz(5):) => fD
        call c_f_pointer(E,   fE,  shp); bE (sz(1):, sz(3):, sz(5):) => fE
        call c_f_pointer(F,   fF,  shp); bF (sz(1):, sz(3):, sz(5):) => fF
        call c_f_pointer(TI,  fTI, shp); bTI(sz(1):, sz(3):, sz(5):) => fTI
        call c_f_pointer(RES, fR,  shp); bR (sz(1):, sz(3):, sz(5):) => fR
        call c_f_pointer(vecw, fw, [hi-lo+1])
        call c_f_pointer(vecg, fg, [hi-lo+1])

        allocate(th)
        allocate(th%COEF3)
        th%COEF3%D => bD
        th%COEF3%E => bE
        th%COEF3%F => bF
        th%TI      => bTI
        allocate(th%VEC(1))
        allocate(th%VEC(1)%W(lo:hi), th%VEC(1)%G(lo:hi))
        th%VEC(1)%W(lo:hi) = fw
        th%VEC(1)%G(lo:hi) = fg

        call THOMAS_3D(lim(1), lim(2), lim(3), lim(4), lim(5), lim(6), di, dj, th, bR)

        fw = th%VEC(1)%W(lo:hi)
        fg = th%VEC(1)%G(lo:hi)
        deallocate(th%VEC(1)%W, th%VEC(1)%G)
        deallocate(th%VEC)
        deallocate(th%COEF3)
        deallocate(th)
    end subroutine ref_thomas3d

    !> OrlanskiCelerity2D (ModuleFunctions.F90:4129-4490) with the optional
    !> arguments ModuleAdvectionDiffusion passes (AD:5549-5568): the reference
    !> field, FlowVelX and NewValue.  newf/oldf/ref/cp span Size, column-major;
    !> newf(i,j,k) and oldf(i,j,k) are updated as the reference updates them.
    subroutine ref_orlanski2d(sz, newf, oldf, ref, cp, imin, imax, jmin, jmax, di, dj,  &
                              i, j, k, eastnorth, limitmax, dt, flowvelx, newvalue)     &
               bind(c, name='ref_orlanski2d')
        integer(c_int), intent(in)  :: sz(6)
        type(c_ptr), value          :: newf, oldf, ref, cp
        integer(c_int), value       :: imin, imax, jmin, jmax, di, dj, i, j, k, eastnorth
        real(c_double), value       :: limitmax, dt, flowvelx
        real(c_double), intent(out) :: newvalue
        real,    pointer :: fN(:,:,:), fO(:,:,:), fR(:,:,:), bN(:,:,:), bO(:,:,:), bR(:,:,:)
        integer, pointer :: fC(:,:,:), bC(:,:,:)
        real    :: nv
        integer :: shp(3)

        shp = [sz(2)-sz(1)+1, sz(4)-sz(3)+1, sz(6)-sz(5)+1]
        call c_f_pointer(newf, fN, shp); bN(sz(1):, sz(3):, sz(5):) => fN
        call c_f_pointer(oldf, fO, shp); bO(sz(1):, sz(3):, sz(5):) => fO
        call c_f_pointer(ref,  fR, shp); bR(sz(1):, sz(3):, sz(5):) => fR
        call c_f_pointer(cp,   fC, shp); bC(sz(1):, sz(3):, sz(5):) => fC
        call OrlanskiCelerity2D(NewField = bN, OldField = bO, ComputePoints = bC,       &
                                ReferenceField = bR, Imin = imin, Imax = imax,          &
                                Jmin = jmin, Jmax = jmax, di = di, dj = dj, i = i,      &
                                j = j, k = k, LimitMax = limitmax,                      &
                                EastNorthBoundary = eastnorth /= 0, DT = dt,            &
                                FlowVelX = flowvelx, NewValue = nv)
        newvalue = nv
    end subroutine ref_orlanski2d

    !> ComputeAdvection1D_V2 (ModuleFunctions.F90:10534-10591) on 1-based lines of length n.
    subroutine ref_advection1d_v2(n, ilb, iub, dt, du, prop, q, v, cp, C, D, E, F,  &
                                  method, tvd, volrelmax, upwind2)                  &
               bind(c, name='ref_advection1d_v2')
        integer(c_int), value :: n, ilb, iub, method, tvd, upwind2
        real(c_double), value :: dt, volrelmax
        type(c_ptr), value    :: du, prop, q, v, cp, C, D, E, F
        real,    pointer :: fdu(:), fprop(:), fC(:), fD(:), fE(:), fF(:)
        real(8), pointer :: fq(:), fv(:)
        integer, pointer :: fcp(:)
        real    :: dt_, vr_
        logical :: up2

        call c_f_pointer(du, fdu, [n]);   call c_f_pointer(prop, fprop, [n])
        call c_f_pointer(q, fq, [n]);     call c_f_pointer(v, fv, [n])
        call c_f_pointer(cp, fcp, [n])
        call c_f_pointer(C, fC, [n]);     call c_f_pointer(D, fD, [n])
        call c_f_pointer(E, fE, [n]);     call c_f_pointer(F, fF, [n])
        dt_ = dt; vr_ = volrelmax; up2 = (upwind2 /= 0)
        call ComputeAdvection1D_V2(ilb, iub, dt_, fdu, fprop, fq, fv, fcp,             &
                                   fC, fD, fE, fF, method, tvd, vr_, up2)
    end subroutine ref_advection1d_v2

    !> ComputeAdvection1D_TVD_SuperBee_2 (ModuleFunctions.F90:10642-10699).
    subroutine ref_advection1d_tvd_superbee_2(n, ilb, iub, dt, du, prop, q, v, cp, D, E) &
               bind(c, name='ref_advection1d_tvd_superbee_2')
        integer(c_int), value :: n, ilb, iub
        real(c_double), value :: dt
        type(c_ptr), value    :: du, prop, q, v, cp, D, E
        real,    pointer :: fdu(:), fprop(:), fD(:), fE(:)
        real(8), pointer :: fq(:), fv(:)
        integer, pointer :: fcp(:)
        real :: dt_

        call c_f_pointer(du, fdu, [n]);   call c_f_pointer(prop, fprop, [n])
        call c_f_pointer(q, fq, [n]);     call c_f_pointer(v, fv, [n])
        call c_f_pointer(cp, fcp, [n])
        call c_f_pointer(D, fD, [n]);     call c_f_pointer(E, fE, [n])
        dt_ = dt
        call ComputeAdvection1D_TVD_SuperBee_2(ilb, iub, dt_, fdu, fprop, fq, fv, fcp, fD, fE)
    end subroutine ref_advection1d_tvd_superbee_2

    !> ComputeAdvectionFace (ModuleFunctions.F90:10702-10894) for one face.
    subroutine ref_advection_face(prop4, v4, du4, dt, qface, volrelmax, method, tvd, &
                                  nearb, upwind2, cface)                             &
               bind(c, name='ref_advection_face')
        real(c_double), intent(in)  :: prop4(4), v4(4), du4(4)
        real(c_double), value       :: dt, qface, volrelmax
        integer(c_int), value       :: method, tvd, nearb, upwind2
        real(c_double), intent(out) :: cface(4)
        real    :: p_(4), du_(4), cf(4), dt_, vr_
        real(8) :: v_(4), q_

        p_ = prop4; du_ = du4; v_ = v4; q_ = qface; dt_ = dt; vr_ = volrelmax
        call ComputeAdvectionFace(p_, v_, du_, dt_, q_, vr_, method, tvd,            &
                                  nearb /= 0, upwind2 /= 0, cf)
        cface = cf
    end subroutine ref_advection_face

end module ref_mf_bind
