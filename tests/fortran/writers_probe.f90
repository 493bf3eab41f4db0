! Test driver of the host writers of vpm_lib (writeline, writesol, writesolXavatar;
! vpm.f90:662-833): fills the public grid state with known values and calls the writers,
! without any GPU call.  tests/test_writers.py compiles it, runs it and reads the files back.
program writers_probe
    use pmgrid
    use vpm_vars
    use vpm_lib
    implicit none
    integer, parameter :: NX = 110, NY = 6, NZ = 5
    integer :: i, j, k
    NXpm = NX; NYpm = NY; NZpm = NZ
    XMIN_pm = -360.d0; YMIN_pm = -20.d0; ZMIN_pm = -16.d0
    DXpm = 8.d0; DYpm = 7.d0; DZpm = 6.5d0
    NXs_bl(1) = 2; NXf_bl(1) = NX - 1
    NYs_bl(1) = 2; NYf_bl(1) = NY - 1
    NZs_bl(1) = 1; NZf_bl(1) = NZ
    allocate(velvrx_pm(NX, NY, NZ), velvry_pm(NX, NY, NZ), velvrz_pm(NX, NY, NZ), RHS_pm(4, NX, NY, NZ))
    do k = 1, NZ
        do j = 1, NY
            do i = 1, NX
                velvrx_pm(i, j, k) = 1.d0 * i + 1.d-2 * j + 1.d-4 * k
                velvry_pm(i, j, k) = 2.d0 * i + 2.d-2 * j + 2.d-4 * k
                velvrz_pm(i, j, k) = 3.d0 * i + 3.d-2 * j + 3.d-4 * k
                RHS_pm(1, i, j, k) = 5.d0 * i - 1.d-2 * j
                RHS_pm(2, i, j, k) = 6.d0 * i - 1.d-2 * k
                RHS_pm(3, i, j, k) = 7.d0 * i + 1.d-3 * j * k
                RHS_pm(4, i, j, k) = 1.d0
            enddo
        enddo
    enddo
    NTIME_pm = 1
    call writeline
    NTIME_pm = 2
    call writeline
    NTIME_pm = 7
    call writesol
    call writesolXavatar(1)
    NTIME_pm = 8
    call writesolXavatar(0)
end program writers_probe
