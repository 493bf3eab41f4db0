!-------------------------------------------------------------------------------
! Example driver of the drop-in vpm_lib: the application loop of the reference's
! test.f90 (read particles.bin, vpm(0), remesh, then per step vpm(2), explicit Euler
! x += (u+UINF) dt, QP(1:3) -= GP dt, vpm(0), remesh) without its inflow/outflow logic.
!
!   ring_driver [nsteps] [dt]
!
! Reads pm.inp and particles.bin (test.f90:24-36 format) from the working directory;
! writes first_call.bin (UP, GP of the first vpm(2)) and final.bin (XPR, QPR at the end)
! as stream files: int32 NVR, then the arrays in Fortran order.
!-------------------------------------------------------------------------------
program ring_driver
    use vpm_lib
    use test_mod
    use pmgrid
    implicit none
    double precision :: Vref, NI_in, DT_in, UINF(3)
    integer :: i, j, neq, nsteps
    character(len=64) :: arg

    neq = 3
    nsteps = 1
    DT_in = 0.01d0
    NI_in = -0.1d0
    if (command_argument_count() >= 1) then
        call get_command_argument(1, arg)
        read(arg, *) nsteps
    endif
    if (command_argument_count() >= 2) then
        call get_command_argument(2, arg)
        read(arg, *) DT_in
    endif

    open(1, file='particles.bin', form='unformatted')
    read(1) NVR_ext
    read(1) Vref
    allocate(XPR(3, NVR_ext), QPR(neq + 1, NVR_ext))
    QPR = 0; XPR = 0
    do i = 1, NVR_ext
        read(1) XPR(1, i), XPR(2, i), XPR(3, i), QPR(1, i), QPR(2, i), QPR(3, i)
    enddo
    close(1)
    QPR(1:3, :) = -QPR(1:3, :) * Vref
    QPR(neq + 1, :) = Vref
    mrem = 1
    UINF = 0.d0

    allocate(UPR(3, NVR_ext), GPR(3, NVR_ext))
    call vpm(XPR, QPR, UPR, GPR, NVR_ext, neq, 0, RHS_pm_in, velx, vely, velz, 0, NI_in, NVR_ext)
    deallocate(UPR, GPR)
    call remesh_particles_3d(1)

    do i = 1, nsteps
        allocate(UPR(3, NVR_ext), GPR(3, NVR_ext))
        UPR = 0; GPR = 0
        call vpm(XPR, QPR, UPR, GPR, NVR_ext, neq, 2, RHS_pm_in, velx, vely, velz, i, NI_in, NVR_ext)
        if (i == 1) then
            open(2, file='first_call.bin', access='stream', form='unformatted', status='replace')
            write(2) NVR_ext, UPR, GPR
            close(2)
        endif
        do j = 1, NVR_ext
            XPR(1:3, j) = XPR(1:3, j) + (UPR(1:3, j) + UINF(1:3)) * DT_in
            QPR(1:3, j) = QPR(1:3, j) - GPR(1:3, j) * DT_in
        enddo
        call vpm(XPR, QPR, UPR, GPR, NVR_ext, neq, 0, RHS_pm_in, velx, vely, velz, i, NI_in, NVR_ext)
        deallocate(UPR, GPR)
        call remesh_particles_3d(1)
    enddo

    open(2, file='final.bin', access='stream', form='unformatted', status='replace')
    write(2) NVR_ext, XPR, QPR
    close(2)
    write(*,*) 'ring_driver done', nsteps, NVR_ext
end program ring_driver
