!-------------------------------------------------------------------------------
! Drop-in Fortran front end of the MI355X VPM core.
!
! Same module names, entry points and call modes as the reference library
! (papis84/vpm_free): vpm(XP,QP,UP,GP,NVR,neqpm,WhatToDo,RHS_pm,velx,vely,velz,NTIME,NI,NVRM)
! (vpm.f90:54-74), the external remesh_particles_3d(iflag) (vpm_remesh.f90:5-225), the
! public state modules pmgrid/vpm_vars/vpm_size/test_mod/projlib/openmpth, and the pm.inp
! file read at every WhatToDo=0 (vpm.f90:132-152).  The host keeps the writers
! (writesol, writeline, writesolXavatar; vpm.f90:662-833).
!
! All numerical work is done by libvpmx.so on one GPU through the C ABI of
! include/vpmx.h; the MPI ranks of the reference collapse into that device.
!-------------------------------------------------------------------------------
module vpmx_bind
    use iso_c_binding
    implicit none
    integer, parameter :: VPMX_MAXB = 64

    type, bind(C) :: vpmx_params_t
        real(c_double)     :: dpm(3)
        integer(c_int32_t) :: interf_iproj, ibctyp, idvpm
        real(c_double)     :: epsvol
        integer(c_int32_t) :: ncoarse, nbi, nbj, nbk, nremesh, ncell_rem, iyntree, ilevmax
        integer(c_int32_t) :: ompthreads, idefine, iynslice, ipmwrite
        integer(c_int32_t) :: ipmwstart(10), ipmwsteps(10)
    end type

    type, bind(C) :: vpmx_geom_t
        real(c_double)     :: xbound(6), dpm(3)
        integer(c_int32_t) :: nn(3), nn_bl(6)
        integer(c_int32_t) :: nblocks, nbi, nbj, nbk
        real(c_double)     :: xbound_bl(6*VPMX_MAXB)
        integer(c_int32_t) :: nnbl(3*VPMX_MAXB), nnbl_bl(6*VPMX_MAXB)
        real(c_double)     :: xbound_coarse(6), dpm_coarse(3)
        integer(c_int32_t) :: nn_coarse(3), nn_bl_coarse(6)
        integer(c_int32_t) :: ilevmax, defined
    end type

    interface
        integer(c_int32_t) function vpmx_read_pm_inp(path, p) bind(C, name='vpmx_read_pm_inp')
            import :: c_int32_t, c_char, vpmx_params_t
            character(kind=c_char), intent(in) :: path(*)
            type(vpmx_params_t), intent(out) :: p
        end function
        integer(c_int32_t) function vpmx_create(ctx, device) bind(C, name='vpmx_create')
            import :: c_int32_t, c_ptr
            type(c_ptr), intent(out) :: ctx
            integer(c_int32_t), value :: device
        end function
        function vpmx_last_error(ctx) bind(C, name='vpmx_last_error')
            import :: c_ptr
            type(c_ptr), value :: ctx
            type(c_ptr) :: vpmx_last_error
        end function
        integer(c_int32_t) function vpmx_set_params(ctx, p) bind(C, name='vpmx_set_params')
            import :: c_int32_t, c_ptr, vpmx_params_t
            type(c_ptr), value :: ctx
            type(vpmx_params_t), intent(in) :: p
        end function
        integer(c_int32_t) function vpmx_get_geometry(ctx, g) bind(C, name='vpmx_get_geometry')
            import :: c_int32_t, c_ptr, vpmx_geom_t
            type(c_ptr), value :: ctx
            type(vpmx_geom_t), intent(out) :: g
        end function
        integer(c_int32_t) function vpmx_vpm(ctx, xp, qp, up, gp, nvr, neqpm, whattodo, ntime, ni) &
                bind(C, name='vpmx_vpm')
            import :: c_int32_t, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx, xp, qp, up, gp
            integer(c_int64_t), value :: nvr
            integer(c_int32_t), value :: neqpm, whattodo, ntime
            real(c_double), value :: ni
        end function
        integer(c_int32_t) function vpmx_remesh_count(ctx, xp, qp, nvr, neqpm, mrem, nvr_out) &
                bind(C, name='vpmx_remesh_count')
            import :: c_int32_t, c_int64_t, c_ptr
            type(c_ptr), value :: ctx, xp, qp
            integer(c_int64_t), value :: nvr
            integer(c_int32_t), value :: neqpm, mrem
            integer(c_int64_t), intent(out) :: nvr_out
        end function
        integer(c_int32_t) function vpmx_remesh_fetch(ctx, xp, qp) bind(C, name='vpmx_remesh_fetch')
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: ctx, xp, qp
        end function
        integer(c_int32_t) function vpmx_get_grid(ctx, which, dst, cap) bind(C, name='vpmx_get_grid')
            import :: c_int32_t, c_int64_t, c_ptr
            type(c_ptr), value :: ctx, dst
            integer(c_int32_t), value :: which
            integer(c_int64_t), value :: cap
        end function
        function c_strlen(s) bind(C, name='strlen')
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
            integer(c_size_t) :: c_strlen
        end function
    end interface

    type(c_ptr), save :: vpmx_ctx_h = c_null_ptr

contains

    subroutine vpmx_check(st, where)
        integer(c_int32_t), intent(in) :: st
        character(*), intent(in) :: where
        character(len=512) :: msg
        character(kind=c_char), pointer :: chars(:)
        type(c_ptr) :: p
        integer :: n, i
        if (st == 0) return
        msg = ' '
        if (c_associated(vpmx_ctx_h)) then
            p = vpmx_last_error(vpmx_ctx_h)
            n = int(min(c_strlen(p), int(len(msg), c_size_t)))
            call c_f_pointer(p, chars, [n])
            do i = 1, n
                msg(i:i) = chars(i)
            enddo
        endif
        write(*,*) 'vpm_lib: ', where, ' failed, status ', st, ': ', trim(msg)
        error stop 'vpm_lib: libvpmx call failed'
    end subroutine vpmx_check

    subroutine vpmx_ensure_ctx()
        integer(c_int32_t) :: st
        if (c_associated(vpmx_ctx_h)) return
        st = vpmx_create(vpmx_ctx_h, 0_c_int32_t)
        if (st /= 0) then
            write(*,*) 'vpm_lib: no usable GPU (vpmx_create status ', st, ')'
            error stop 'vpm_lib: vpmx_create failed'
        endif
    end subroutine vpmx_ensure_ctx
end module vpmx_bind

!------------------------------------------------------------------ public state
module pmgrid
    double precision, allocatable, target, save :: velvrx_pm(:,:,:), velvry_pm(:,:,:), velvrz_pm(:,:,:)
    double precision, allocatable, target, save :: RHS_pm(:,:,:,:)
    double precision, save :: XMIN_pm, XMAX_pm, YMIN_pm, YMAX_pm, ZMIN_pm, ZMAX_pm
    double precision, save :: DD, DXpm, DYpm, DZpm, DXpm2, DYpm2, DZpm2, DVpm, EPSVOL
    integer, save :: NXpm, NYpm, NZpm, NXpm_par, NYpm_par, NZpm_par
    integer, save :: NXs_bl(10), NYs_bl(10), NXf_bl(10), NYf_bl(10), NZs_bl(10), NZf_bl(10), NBlocks
end module pmgrid

module pmeshpar
    double precision, save :: PI, PI2, PI4, DT
    integer, save :: nbound, ndumcell, NVR_CFD_sa, IDVPM
    integer :: levmax, npar_cell, ND
end module pmeshpar

module parvar
    integer :: NVR
    double precision, pointer, save :: XP(:,:), QP(:,:)
    double precision, pointer, save :: UP(:,:), GP(:,:)
end module parvar

module test_mod
    double precision, allocatable, target :: XPR(:,:), QPR(:,:), UPR(:,:), GPR(:,:), XPO(:,:), QPO(:,:), &
                                             XP_in(:,:), QP_in(:,:), XSOUR(:,:), QSOUR(:,:), XP_all(:,:), QP_all(:,:)
    double precision, pointer :: velx(:,:,:), vely(:,:,:), velz(:,:,:)
    double precision, pointer :: RHS_pm_in(:,:,:,:), RHS_pm_out(:,:,:,:)
    integer, allocatable :: qflag(:)
    integer :: NVR_ext, NVR_sources, NVR_sources_init, NVR_ext_init
end module test_mod

module vpm_vars
    double precision :: DT_c, V_ref, NI
    integer :: interf_iproj, ncell_rem, iynhj
    integer :: ncoarse, nparcell1d
    integer :: neqpm, NVR_p, NVR_size, iwrite, NTIME_pm
    integer, save :: IPMWRITE, mrem = 1, idefine, iynslice
    integer, save :: IPMWSTART(10), IPMWSTEPS(10)
end module vpm_vars

module vpm_size
    double precision, save :: Xbound(6), Dpm(3), Xbound0(6), Dpm0(3)
    integer, save :: NN_bl(6), NN(3), NN0_bl(6), NN0(3)
    double precision, allocatable, save :: Xbound_bl(:,:)
    integer, allocatable, save :: NNbl_bl(:,:), NNbl(:,:)
    double precision, save :: Xbound_tmp(6), Xbound_coarse(6), Dpm_coarse(3)
    integer, save :: NN_tmp(3), NN_bl_tmp(6), NN_coarse(3), NN_bl_coarse(6)
    integer, save :: nb_i, nb_j, nb_k, NBB, NXbl, NYbl, NZbl, BLOCKS, NXB, NYB, NZB, ndumcell_coarse, ndumcell_bl
    double precision :: starttime, endtime, st, et, ct
    integer, save :: iynbc, iret, NBI, NBJ, NBK, NVR_out_thres, NREMESH, ntorder, &
                     iyntree, ilevmax, itree, nsize_out(3), ibctyp, NWRITE
end module vpm_size

module openmpth
    integer :: OMPTHREADS
end module openmpth

! The reference's test driver `use`s projlib; its projection function is available here.
module projlib
contains
    double precision function proj_fun(itype, x)
        integer, intent(in) :: itype
        double precision, intent(in) :: x
        double precision :: xabs
        xabs = abs(x)
        proj_fun = 0.d0
        if (itype == 4) then
            if (xabs < 1.d0) then
                proj_fun = 1.d0 - 2.5d0 * xabs**2 + 1.5d0 * xabs**3
            else if (xabs <= 2.d0) then
                proj_fun = 0.5d0 * (2.d0 - xabs)**2 * (1.d0 - xabs)
            endif
        else if (itype == 3) then
            if (xabs < 0.5d0) then
                proj_fun = 0.5d0 * (xabs + 1.5d0)**2 - 1.5d0 * (xabs + 0.5d0)**2
            else if (xabs <= 1.5d0) then
                proj_fun = 0.5d0 * (-xabs + 1.5d0)**2
            endif
        else if (itype == 2) then
            if (xabs <= 1.d0) proj_fun = 1.d0 - xabs
        endif
    end function proj_fun
end module projlib

!------------------------------------------------------------------ vpm_lib
module vpm_lib
    use vpm_vars
    use vpm_size
    use openmpth
    implicit none
contains

    subroutine vpm(XP_in, QP_in, UP_in, GP_in, NVR_in, neqpm_in, WhatToDo, &
                   RHS_pm_in, Velx, Vely, Velz, NTIME_in, NI_in, NVRM_in)
        use iso_c_binding
        use vpmx_bind
        use pmgrid
        use pmeshpar
        implicit none
        ! contiguous: the C ABI takes the arrays' base addresses (whole arrays in practice)
        double precision, intent(inout), target, contiguous :: XP_in(:,:), QP_in(:,:), UP_in(:,:), GP_in(:,:)
        double precision, intent(inout), pointer :: RHS_pm_in(:,:,:,:)
        double precision, intent(inout), pointer :: velx(:,:,:), vely(:,:,:), velz(:,:,:)
        integer, intent(inout) :: NVR_in
        integer, intent(in) :: neqpm_in, WhatToDo, NVRM_in, NTIME_in
        double precision, intent(in) :: NI_in
        type(vpmx_params_t) :: prm
        integer(c_int32_t) :: st
        type(c_ptr) :: pu, pg
        integer :: i, k, j

        NVR_size = NVRM_in
        NTIME_pm = NTIME_in
        neqpm = neqpm_in
        NI = NI_in
        PI = 4.d0 * atan(1.d0)
        PI2 = 2.d0 * PI
        PI4 = 4.d0 * PI
        nullify(RHS_pm_in)
        call vpmx_ensure_ctx()

        if (WhatToDo == 0) then
            ND = 3
            st = vpmx_read_pm_inp('pm.inp' // c_null_char, prm)
            if (st /= 0) then
                write(*,*) 'vpm_lib: cannot read pm.inp (status ', st, ')'
                error stop 'vpm_lib: pm.inp'
            endif
            call vpmx_check(vpmx_set_params(vpmx_ctx_h, prm), 'set_params')
            call copy_params(prm)
            call vpmx_check(vpmx_vpm(vpmx_ctx_h, c_loc(XP_in), c_loc(QP_in), c_null_ptr, c_null_ptr, &
                            int(NVR_in, c_int64_t), int(neqpm, c_int32_t), 0_c_int32_t, &
                            int(NTIME_pm, c_int32_t), NI), 'define_sizes')
            call copy_geometry()
            if (allocated(velvrx_pm)) deallocate(velvrx_pm, velvry_pm, velvrz_pm)
            allocate(velvrx_pm(NXpm,NYpm,NZpm), velvry_pm(NXpm,NYpm,NZpm), velvrz_pm(NXpm,NYpm,NZpm))
            velvrx_pm = 0.d0; velvry_pm = 0.d0; velvrz_pm = 0.d0
            nullify(velx); nullify(vely); nullify(velz)
            velx => velvrx_pm; vely => velvry_pm; velz => velvrz_pm
            iwrite = 0
            return
        endif

        if (NVR_in == 0) return   ! vpm.f90:171-172
        if (WhatToDo == 1 .or. WhatToDo == 2 .or. WhatToDo == 3) then
            pu = c_null_ptr; pg = c_null_ptr
            if (WhatToDo == 2) then
                pu = c_loc(UP_in); pg = c_loc(GP_in)
            endif
            call vpmx_check(vpmx_vpm(vpmx_ctx_h, c_loc(XP_in), c_loc(QP_in), pu, pg, &
                            int(NVR_in, c_int64_t), int(neqpm, c_int32_t), int(WhatToDo, c_int32_t), &
                            int(NTIME_pm, c_int32_t), NI), 'vpm')
            if (WhatToDo == 1 .or. WhatToDo == 2) then
                if (allocated(velvrx_pm)) deallocate(velvrx_pm, velvry_pm, velvrz_pm)
                allocate(velvrx_pm(NXpm,NYpm,NZpm), velvry_pm(NXpm,NYpm,NZpm), velvrz_pm(NXpm,NYpm,NZpm))
                call fetch_velocity()
            endif
            if (WhatToDo == 1) then
                nullify(velx); nullify(vely); nullify(velz)
                velx => velvrx_pm; vely => velvry_pm; velz => velvrz_pm
            endif
            if (WhatToDo == 2) then
                ! host outputs of the reference's mode 2 (vpm.f90:275-284)
                if (IPMWRITE > 0) then
                    do i = 1, IPMWRITE
                        if (NTIME_pm >= IPMWSTART(i) .and. NTIME_pm <= (IPMWSTART(i) + IPMWSTEPS(i))) then
                            call fetch_rhs()
                            call writesol
                        endif
                    enddo
                endif
                if (iynslice /= 0) then
                    call fetch_rhs()
                    call writesolXavatar(iynslice)
                endif
                call writeline
                k = int(0.5 * NZpm); j = int(0.5 * NYpm)
                if (iynslice == -1) then
                    do i = NXs_bl(1), NXf_bl(1)
                        write(32,*) velvrx_pm(i,j,k), velvry_pm(i,j,k), velvrz_pm(i,j,k)
                    enddo
                endif
                iwrite = 0
                ! mode 2 frees the grids; velx/vely/velz dangle as in the reference (vpm.f90:308-311)
                deallocate(velvrx_pm, velvry_pm, velvrz_pm)
                if (allocated(RHS_pm)) deallocate(RHS_pm)
            endif
            return
        endif
        if (WhatToDo == 4 .or. WhatToDo == 5) then
            ! 4: stretching from the velocity grids of the last 0/1/5 call (vpm.f90:316-343);
            ! 5: diffusion, UP = 0, GP = interpolated -NI lap(w) * QP(4) (vpm.f90:345-373)
            call vpmx_check(vpmx_vpm(vpmx_ctx_h, c_loc(XP_in), c_loc(QP_in), c_loc(UP_in), c_loc(GP_in), &
                            int(NVR_in, c_int64_t), int(neqpm, c_int32_t), int(WhatToDo, c_int32_t), &
                            int(NTIME_pm, c_int32_t), NI), 'vpm')
            if (WhatToDo == 4) then
                if (IPMWRITE > 0) then
                    do i = 1, IPMWRITE
                        if (NTIME_pm >= IPMWSTART(i) .and. NTIME_pm <= (IPMWSTART(i) + IPMWSTEPS(i))) then
                            call fetch_rhs()
                            call writesol
                        endif
                    enddo
                endif
                if (iynslice /= 0) then
                    call fetch_rhs()
                    call writesolXavatar(iynslice)
                endif
            else
                if (allocated(velvrx_pm)) deallocate(velvrx_pm, velvry_pm, velvrz_pm)
                allocate(velvrx_pm(NXpm,NYpm,NZpm), velvry_pm(NXpm,NYpm,NZpm), velvrz_pm(NXpm,NYpm,NZpm))
                velvrx_pm = 0.d0; velvry_pm = 0.d0; velvrz_pm = 0.d0
                nullify(velx); nullify(vely); nullify(velz)
                velx => velvrx_pm; vely => velvry_pm; velz => velvrz_pm
            endif
            iwrite = 0
            if (allocated(RHS_pm)) deallocate(RHS_pm)
            return
        endif
        write(*,*) 'vpm_lib: WhatToDo = ', WhatToDo, ' is not supported by this build'
        error stop 'vpm_lib: WhatToDo'
    end subroutine vpm

    subroutine copy_params(p)
        use vpmx_bind
        use pmgrid
        use pmeshpar
        type(vpmx_params_t), intent(in) :: p
        DXpm = p%dpm(1); DYpm = p%dpm(2); DZpm = p%dpm(3)
        interf_iproj = p%interf_iproj
        ibctyp = p%ibctyp
        IDVPM = p%idvpm
        EPSVOL = p%epsvol
        ncoarse = p%ncoarse
        NBI = p%nbi; NBJ = p%nbj; NBK = p%nbk
        NREMESH = p%nremesh; ncell_rem = p%ncell_rem
        iyntree = p%iyntree; ilevmax = p%ilevmax
        OMPTHREADS = p%ompthreads
        idefine = p%idefine
        iynslice = p%iynslice
        IPMWRITE = p%ipmwrite
        IPMWSTART = p%ipmwstart
        IPMWSTEPS = p%ipmwsteps
    end subroutine copy_params

    ! Module state of define_sizes (vpm.f90:539-546, 556-612) from the library's geometry.
    subroutine copy_geometry()
        use vpmx_bind
        use pmgrid
        type(vpmx_geom_t) :: g
        integer :: b
        call vpmx_check(vpmx_get_geometry(vpmx_ctx_h, g), 'get_geometry')
        Xbound = g%xbound; Dpm = g%dpm; NN = g%nn; NN_bl = g%nn_bl
        XMIN_pm = Xbound(1); YMIN_pm = Xbound(2); ZMIN_pm = Xbound(3)
        XMAX_pm = Xbound(4); YMAX_pm = Xbound(5); ZMAX_pm = Xbound(6)
        NXpm = NN(1); NYpm = NN(2); NZpm = NN(3)
        NXs_bl(1) = NN_bl(1); NYs_bl(1) = NN_bl(2); NZs_bl(1) = NN_bl(3)
        NXf_bl(1) = NN_bl(4); NYf_bl(1) = NN_bl(5); NZf_bl(1) = NN_bl(6)
        DXpm = Dpm(1); DYpm = Dpm(2); DZpm = Dpm(3)
        DVpm = DXpm * DYpm * DZpm
        DXpm2 = 2 * DXpm; DYpm2 = 2 * DYpm; DZpm2 = 2 * DZpm
        BLOCKS = g%nblocks
        if (allocated(Xbound_bl)) deallocate(Xbound_bl, NNbl, NNbl_bl)
        allocate(Xbound_bl(6, BLOCKS), NNbl(3, BLOCKS), NNbl_bl(6, BLOCKS))
        do b = 1, BLOCKS
            Xbound_bl(:, b) = g%xbound_bl(6*(b-1)+1:6*b)
            NNbl(:, b) = g%nnbl(3*(b-1)+1:3*b)
            NNbl_bl(:, b) = g%nnbl_bl(6*(b-1)+1:6*b)
        enddo
        nb_i = 1; nb_j = 1; nb_k = 1
        Xbound_coarse = g%xbound_coarse; Dpm_coarse = g%dpm_coarse
        NN_coarse = g%nn_coarse; NN_bl_coarse = g%nn_bl_coarse
        ndumcell_bl = ncoarse; ndumcell_coarse = 4
        ilevmax = g%ilevmax
        NXB = nint((Xbound(4) - Xbound(1)) / Dpm(1))
        NYB = nint((Xbound(5) - Xbound(2)) / Dpm(2))
        NZB = nint((Xbound(6) - Xbound(3)) / Dpm(3))
        NXbl = NXB / NBI; NYbl = NYB / NBJ; NZbl = NZB / NBK
    end subroutine copy_geometry

    subroutine fetch_velocity()
        use vpmx_bind
        use pmgrid
        double precision, allocatable, target :: buf(:)
        integer(c_int64_t) :: n
        n = int(NXpm, c_int64_t) * NYpm * NZpm
        allocate(buf(3 * n))
        call vpmx_check(vpmx_get_grid(vpmx_ctx_h, 3_c_int32_t, c_loc(buf), 3 * n), 'get_grid(vel)')
        velvrx_pm = reshape(buf(1:n), [NXpm, NYpm, NZpm])
        velvry_pm = reshape(buf(n+1:2*n), [NXpm, NYpm, NZpm])
        velvrz_pm = reshape(buf(2*n+1:3*n), [NXpm, NYpm, NZpm])
        deallocate(buf)
    end subroutine fetch_velocity

    subroutine fetch_rhs()
        use vpmx_bind
        use pmgrid
        integer(c_int64_t) :: n
        if (allocated(RHS_pm)) then
            if (size(RHS_pm, 2) == NXpm .and. size(RHS_pm, 3) == NYpm .and. size(RHS_pm, 4) == NZpm) return
            deallocate(RHS_pm)
        endif
        allocate(RHS_pm(neqpm + 1, NXpm, NYpm, NZpm))
        n = int(neqpm + 1, c_int64_t) * NXpm * NYpm * NZpm
        call vpmx_check(vpmx_get_grid(vpmx_ctx_h, 0_c_int32_t, c_loc(RHS_pm), n), 'get_grid(rhs)')
    end subroutine fetch_rhs

    ! writesol (vpm.f90:662-709): Tecplot-style unformatted point file of the inner grid.
    subroutine writesol
        use pmgrid
        character(len=50) :: filout
        integer :: i, j, k
        double precision :: XPM, YPM, ZPM
        write(filout, '(i5.5,a)') NTIME_pm, 'solution.bin'
        open(1, file=filout, form='unformatted')
        write(1) 'VARIABLES = "X" "Y" "Z" "U" "V" "W" "VORTX" "VORTY" "VORTZ"'
        write(1) 'ZONE I=', NXf_bl(1) - NXs_bl(1) + 1, ' J=', NYf_bl(1) - NYs_bl(1) + 1, &
                 ' K=', NZf_bl(1) - NZs_bl(1) + 1, ' F=POINT'
        do k = NZs_bl(1), NZf_bl(1)
            do j = NYs_bl(1), NYf_bl(1)
                do i = NXs_bl(1), NXf_bl(1)
                    XPM = XMIN_pm + (i - 1) * DXpm
                    YPM = YMIN_pm + (j - 1) * DYpm
                    ZPM = ZMIN_pm + (k - 1) * DZpm
                    write(1) XPM, YPM, ZPM, velvrx_pm(i,j,k), velvry_pm(i,j,k), velvrz_pm(i,j,k), &
                             -RHS_pm(1,i,j,k), -RHS_pm(2,i,j,k), -RHS_pm(3,i,j,k)
                enddo
            enddo
        enddo
        close(1)
        iwrite = 1
    end subroutine writesol

    ! writesolXavatar (vpm.f90:711-787).  The reference tests an undeclared (implicitly typed,
    ! never set) variable to pick the slice positions; the argument is used here.
    subroutine writesolXavatar(itypesliceavatar)
        use pmgrid
        integer, intent(in) :: itypesliceavatar
        character(len=50) :: filout
        integer :: i, j, k, ii, NX_AVA(15), NXPOS
        double precision :: XPM, YPM, ZPM, POSX(15)
        NXPOS = 10
        if (itypesliceavatar == 1) then
            POSX(1:9) = [-350.d0, -250.d0, -150.d0, -50.d0, 50.d0, 150.d0, 250.d0, 350.d0, 450.d0]
            NX_AVA(1) = 2
            NX_AVA(2:10) = int((POSX(1:9) - XMIN_pm) / DXpm) + 1
        else
            POSX(1:9) = [-204.d0, -102.d0, -50.d0, -4.d0, 4.d0, 50.d0, 102.d0, 204.d0, 306.d0]
            NX_AVA(1) = 2
            NX_AVA(2:10) = int((POSX(1:9) - 7.1 - XMIN_pm) / DXpm) + 1
        endif
        write(filout, '(i5.5,a)') NTIME_pm, 'solX.bin'
        open(1, file=filout, form='unformatted')
        do ii = 1, NXPOS
            i = NX_AVA(ii)
            write(1) NYf_bl(1) - NYs_bl(1) + 1, NZf_bl(1) - NZs_bl(1) + 1
            do k = NZs_bl(1), NZf_bl(1)
                do j = NYs_bl(1), NYf_bl(1)
                    XPM = XMIN_pm + (i - 1) * DXpm
                    YPM = YMIN_pm + (j - 1) * DYpm
                    ZPM = ZMIN_pm + (k - 1) * DZpm
                    write(1) XPM, YPM, ZPM, velvrx_pm(i,j,k), velvry_pm(i,j,k), velvrz_pm(i,j,k), &
                             -RHS_pm(1,i,j,k), -RHS_pm(2,i,j,k), -RHS_pm(3,i,j,k)
                enddo
            enddo
        enddo
        close(1)
    end subroutine writesolXavatar

    ! writeline (vpm.f90:789-833): velocity on 9 x-lines, appended to NNhist.bin every call.
    subroutine writeline
        use pmgrid
        character(len=50) :: filout
        integer :: i, j, k, jmat(9), kmat(9), NNJ, NNK, il
        NNJ = NYf_bl(1) - NYs_bl(1) + 1
        NNK = NZf_bl(1) - NZs_bl(1) + 1
        jmat(1) = 0.5 * NNJ;  kmat(1) = 0.5 * NNK
        jmat(2) = 0.25 * NNJ; kmat(2) = 0.25 * NNK
        jmat(3) = 0.25 * NNJ; kmat(3) = 0.75 * NNK
        jmat(4) = 0.75 * NNJ; kmat(4) = 0.25 * NNK
        jmat(5) = 0.75 * NNJ; kmat(5) = 0.75 * NNK
        jmat(6) = 0.5 * NNJ;  kmat(6) = 0.25 * NNK
        jmat(7) = 0.5 * NNJ;  kmat(7) = 0.75 * NNK
        jmat(8) = 0.25 * NNJ; kmat(8) = 0.5 * NNK
        jmat(9) = 0.75 * NNJ; kmat(9) = 0.5 * NNK
        do i = 1, 9
            j = jmat(i); k = kmat(i)
            write(filout, '(i2.2,a)') i, 'hist.bin'
            open(1, file=filout, access='APPEND', form='unformatted')
            if (NTIME_pm == 1) then
                rewind(1)
                write(1) XMIN_pm, YMIN_pm, ZMIN_pm
                write(1) DXpm, DYpm, DZpm
                write(1) NXs_bl(1), NXf_bl(1), j, k
            endif
            write(1) NTIME_pm, (velvrx_pm(il,j,k), il = NXs_bl(1), NXf_bl(1)), &
                               (velvry_pm(il,j,k), il = NXs_bl(1), NXf_bl(1)), &
                               (velvrz_pm(il,j,k), il = NXs_bl(1), NXf_bl(1))
            close(1)
        enddo
    end subroutine writeline
end module vpm_lib

! remesh_particles_3d(iflag) (vpm_remesh.f90:5-225): regenerates test_mod's XPR/QPR on the
! grid nodes (ncell_rem = 1), keeping nodes with |w| >= 1e-9 in k->j->i order; mrem from
! vpm_vars must be set by the caller.
subroutine remesh_particles_3d(iflag)
    use iso_c_binding
    use vpmx_bind
    use vpm_vars
    use test_mod
    implicit none
    integer, intent(in) :: iflag
    integer(c_int64_t) :: nout
    if (iflag /= 1) then
        write(*,*) 'remesh_particles_3d: only iflag=1 is supported'
        error stop 'remesh_particles_3d'
    endif
    call vpmx_ensure_ctx()
    call vpmx_check(vpmx_remesh_count(vpmx_ctx_h, c_loc(XPR), c_loc(QPR), int(NVR_ext, c_int64_t), &
                    int(neqpm, c_int32_t), int(mrem, c_int32_t), nout), 'remesh_count')
    deallocate(XPR, QPR)
    allocate(XPR(3, nout), QPR(neqpm + 1, nout))
    call vpmx_check(vpmx_remesh_fetch(vpmx_ctx_h, c_loc(XPR), c_loc(QPR)), 'remesh_fetch')
    NVR_ext = int(nout)
    NVR_size = NVR_ext
    write(*,*) 'remesh complete', NVR_ext
end subroutine remesh_particles_3d
