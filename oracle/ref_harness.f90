!! Test-only driver for the compiled reference (oracle/_ref).
!!
!! This file is written for this project; it contains no reference code. It
!! sets up the reference's module state (blk, blk0, flags, constants) from
!! plain arrays so that the reference's own module-state routines --
!! physical_bc, UpdateBuffer, average1/average2, UpdateBufferCoordinate,
!! residual -- and its per-block kernels can be called from Python (ctypes)
!! on synthetic multi-block inputs.  It is built only in the development
!! container by oracle/Makefile; nothing here ships with the product.
module refh
    use iso_c_binding
    implicit none
contains

    !! Index ranges of block m by the rule of sub_init_allocatememory.f90:73-144
    subroutine refh_ranges(ni, nj, nk, nsub, sub, r)
        use flag_var
        integer, intent(in) :: ni, nj, nk, nsub
        integer, intent(in) :: sub(11, nsub)
        integer, intent(out) :: r(18)
        integer :: s
        r = 0
        r(7) = 1; r(8) = ni; r(9) = 1; r(10) = nj; r(11) = 1; r(12) = nk
        r(1:6) = r(7:12)
        do s = 1, nsub
            if (sub(9, s) > 0) then
                select case (sub(2, s))
                case (1); r(1) = r(7) - bufferLength
                case (2); r(2) = r(8) + bufferLength
                case (3); r(3) = r(9) - bufferLength
                case (4); r(4) = r(10) + bufferLength
                case (5); r(5) = r(11) - bufferLength
                case (6); r(6) = r(12) + bufferLength
                end select
            end if
        end do
        r(13) = merge(r(7) + 1, r(7), r(7) == r(1))
        r(14) = merge(r(8) - 1, r(8), r(8) == r(2))
        r(15) = merge(r(9) + 1, r(9), r(9) == r(3))
        r(16) = merge(r(10) - 1, r(10), r(10) == r(4))
        r(17) = merge(r(11) + 1, r(11), r(11) == r(5))
        r(18) = merge(r(12) - 1, r(12), r(12) == r(6))
    end subroutine
end module refh

subroutine refh_begin(nblk, dims, nsubs, subtab) bind(C, name="refh_begin")
    use iso_c_binding
    use refh
    use blk_var
    use glbindex_var
    use index_var
    use mpi_var
    use flag_var
    implicit none
    integer(c_int), value :: nblk
    integer(c_int) :: dims(3, nblk), nsubs(nblk)
    integer(c_int) :: subtab(11, *)
    integer :: mm, s, off, r(18), is, ie, js, je, ks, ke, t
    logical :: inited

    call MPI_INITIALIZED(inited, ierr)
    if (.not. inited) call MPI_INIT(ierr)
    myid = 0; numprocs = 1
    bufferLength = 2
    nblock = nblk; blk_loop = nblk; myn = nblk; lastn = nblk
    if (allocated(blk)) deallocate(blk)
    if (allocated(blk0)) deallocate(blk0)
    allocate(blk(nblk), blk0(nblk))
    totalnodes = 0
    off = 0
    do mm = 1, nblk
        call refh_ranges(dims(1, mm), dims(2, mm), dims(3, mm), nsubs(mm), subtab(:, off + 1:off + nsubs(mm)), r)
        blk(mm)%ni = dims(1, mm); blk(mm)%nj = dims(2, mm); blk(mm)%nk = dims(3, mm)
        blk0(mm)%ni = dims(1, mm); blk0(mm)%nj = dims(2, mm); blk0(mm)%nk = dims(3, mm)
        totalnodes = totalnodes + dims(1, mm) * dims(2, mm) * dims(3, mm)
        blk(mm)%is = r(1); blk(mm)%ie = r(2); blk(mm)%js = r(3); blk(mm)%je = r(4); blk(mm)%ks = r(5); blk(mm)%ke = r(6)
        blk(mm)%is0 = r(7); blk(mm)%ie0 = r(8); blk(mm)%js0 = r(9); blk(mm)%je0 = r(10); blk(mm)%ks0 = r(11); blk(mm)%ke0 = r(12)
        blk(mm)%is1 = r(13); blk(mm)%ie1 = r(14); blk(mm)%js1 = r(15); blk(mm)%je1 = r(16); blk(mm)%ks1 = r(17); blk(mm)%ke1 = r(18)
        blk0(mm)%is = r(1); blk0(mm)%ie = r(2); blk0(mm)%js = r(3); blk0(mm)%je = r(4); blk0(mm)%ks = r(5); blk0(mm)%ke = r(6)
        blk0(mm)%is0 = r(7); blk0(mm)%ie0 = r(8); blk0(mm)%js0 = r(9); blk0(mm)%je0 = r(10); blk0(mm)%ks0 = r(11); blk0(mm)%ke0 = r(12)
        blk0(mm)%is1 = r(13); blk0(mm)%ie1 = r(14); blk0(mm)%js1 = r(15); blk0(mm)%je1 = r(16); blk0(mm)%ks1 = r(17); blk0(mm)%ke1 = r(18)
        blk(mm)%num_subface = nsubs(mm); blk0(mm)%num_subface = nsubs(mm)
        allocate(blk(mm)%bc(nsubs(mm)), blk0(mm)%bc0(nsubs(mm)))
        do s = 1, nsubs(mm)
            t = off + s
            blk(mm)%bc(s)%f_no = subtab(1, t); blk(mm)%bc(s)%face = subtab(2, t)
            blk(mm)%bc(s)%is = subtab(3, t); blk(mm)%bc(s)%ie = subtab(4, t)
            blk(mm)%bc(s)%js = subtab(5, t); blk(mm)%bc(s)%je = subtab(6, t)
            blk(mm)%bc(s)%ks = subtab(7, t); blk(mm)%bc(s)%ke = subtab(8, t)
            blk(mm)%bc(s)%blk_t = subtab(9, t); blk(mm)%bc(s)%subface_t = subtab(10, t)
            blk(mm)%bc(s)%ori = subtab(11, t); blk(mm)%bc(s)%fwh = 0
            blk(mm)%bc(s)%dim_f = (subtab(4, t) - subtab(3, t) + 1) * (subtab(6, t) - subtab(5, t) + 1) &
                                  * (subtab(8, t) - subtab(7, t) + 1)
            blk(mm)%bc(s)%rank_c = 0; blk(mm)%bc(s)%rank_t = 0
            blk0(mm)%bc0(s)%f_no = subtab(1, t); blk0(mm)%bc0(s)%face = subtab(2, t)
            blk0(mm)%bc0(s)%is = subtab(3, t); blk0(mm)%bc0(s)%ie = subtab(4, t)
            blk0(mm)%bc0(s)%js = subtab(5, t); blk0(mm)%bc0(s)%je = subtab(6, t)
            blk0(mm)%bc0(s)%ks = subtab(7, t); blk0(mm)%bc0(s)%ke = subtab(8, t)
            blk0(mm)%bc0(s)%blk_t = subtab(9, t); blk0(mm)%bc0(s)%subface_t = subtab(10, t)
            blk0(mm)%bc0(s)%ori = subtab(11, t); blk0(mm)%bc0(s)%fwh = 0
            blk0(mm)%bc0(s)%dim_f = blk(mm)%bc(s)%dim_f
            blk0(mm)%bc0(s)%rank_c = 0; blk0(mm)%bc0(s)%rank_t = 0
        end do
        off = off + nsubs(mm)
        is = r(1); ie = r(2); js = r(3); je = r(4); ks = r(5); ke = r(6)
        allocate(blk(mm)%x(is:ie, js:je, ks:ke), blk(mm)%y(is:ie, js:je, ks:ke), blk(mm)%z(is:ie, js:je, ks:ke))
        allocate(blk(mm)%q(5, is:ie, js:je, ks:ke), blk(mm)%q0(-1:0, 5, is:ie, js:je, ks:ke))
        allocate(blk(mm)%pri_v(7, is:ie, js:je, ks:ke))
        allocate(blk(mm)%rhsi(5, is:ie, js:je, ks:ke), blk(mm)%rhsv(5, is:ie, js:je, ks:ke), blk(mm)%rhs(5, is:ie, js:je, ks:ke))
        allocate(blk(mm)%dudx(9, is:ie, js:je, ks:ke), blk(mm)%tao(6, is:ie, js:je, ks:ke))
        allocate(blk(mm)%vor(is:ie, js:je, ks:ke), blk(mm)%amu(3, is:ie, js:je, ks:ke), blk(mm)%dt(is:ie, js:je, ks:ke))
        allocate(blk(mm)%dxidx(9, is:ie, js:je, ks:ke), blk(mm)%inv_j(is:ie, js:je, ks:ke))
        allocate(blk(mm)%aalpha(3, is:ie, js:je, ks:ke), blk(mm)%bbeta(3, is:ie, js:je, ks:ke), &
                 blk(mm)%ggamma(3, is:ie, js:je, ks:ke))
        allocate(blk(mm)%dist(is:ie, js:je, ks:ke), blk(mm)%spacing(is:ie, js:je, ks:ke), blk(mm)%length(is:ie, js:je, ks:ke))
        allocate(blk(mm)%spci(is:ie, js:je, ks:ke), blk(mm)%spcj(is:ie, js:je, ks:ke), blk(mm)%spck(is:ie, js:je, ks:ke))
        allocate(blk(mm)%vol(is:ie, js:je, ks:ke))
        allocate(blk(mm)%nut(is:ie, js:je, ks:ke), blk(mm)%sa_rhs(is:ie, js:je, ks:ke))
        blk(mm)%x = 0; blk(mm)%y = 0; blk(mm)%z = 0; blk(mm)%q = 0; blk(mm)%q0 = 0; blk(mm)%pri_v = 0
        blk(mm)%rhsi = 0; blk(mm)%rhsv = 0; blk(mm)%rhs = 0; blk(mm)%dudx = 0; blk(mm)%tao = 0; blk(mm)%vor = 0
        blk(mm)%amu = 0; blk(mm)%dt = 0; blk(mm)%dxidx = 0; blk(mm)%inv_j = 0
        blk(mm)%aalpha = 0; blk(mm)%bbeta = 0; blk(mm)%ggamma = 0
        blk(mm)%dist = 0; blk(mm)%spacing = 0; blk(mm)%length = 0; blk(mm)%nut = 0; blk(mm)%sa_rhs = 0
    end do
    !! interface buffers as in sub_init_readbcinfo.f90:440-475
    do mm = 1, nblk
        do s = 1, blk(mm)%num_subface
            if (blk(mm)%bc(s)%blk_t > 0) then
                allocate(blk(mm)%bc(s)%buffer_send(bufferLength, 9, blk(mm)%bc(s)%is:blk(mm)%bc(s)%ie, &
                         blk(mm)%bc(s)%js:blk(mm)%bc(s)%je, blk(mm)%bc(s)%ks:blk(mm)%bc(s)%ke))
                allocate(blk(mm)%bc(s)%buffer(9, blk(mm)%bc(s)%is:blk(mm)%bc(s)%ie, &
                         blk(mm)%bc(s)%js:blk(mm)%bc(s)%je, blk(mm)%bc(s)%ks:blk(mm)%bc(s)%ke))
                t = blk(mm)%bc(s)%blk_t
                allocate(blk(mm)%bc(s)%buffer_recv(bufferLength, 9, &
                         blk0(t)%bc0(blk(mm)%bc(s)%subface_t)%is:blk0(t)%bc0(blk(mm)%bc(s)%subface_t)%ie, &
                         blk0(t)%bc0(blk(mm)%bc(s)%subface_t)%js:blk0(t)%bc0(blk(mm)%bc(s)%subface_t)%je, &
                         blk0(t)%bc0(blk(mm)%bc(s)%subface_t)%ks:blk0(t)%bc0(blk(mm)%bc(s)%subface_t)%ke))
                blk(mm)%bc(s)%buffer_send = 0.d0
                blk(mm)%bc(s)%buffer_recv = 0.d0
                blk(mm)%bc(s)%buffer = 0.d0
            end if
        end do
    end do
end subroutine

subroutine refh_get_ext(mm, r) bind(C, name="refh_get_ext")
    use iso_c_binding
    use blk_var
    implicit none
    integer(c_int), value :: mm
    integer(c_int) :: r(18)
    r = [blk(mm)%is, blk(mm)%ie, blk(mm)%js, blk(mm)%je, blk(mm)%ks, blk(mm)%ke, &
         blk(mm)%is0, blk(mm)%ie0, blk(mm)%js0, blk(mm)%je0, blk(mm)%ks0, blk(mm)%ke0, &
         blk(mm)%is1, blk(mm)%ie1, blk(mm)%js1, blk(mm)%je1, blk(mm)%ks1, blk(mm)%ke1]
end subroutine

!! flags and constants; the non-dimensional freestream follows
!! sub_init_flowfield.f90:30-39
subroutine refh_setup(dimen, solver, itime, turb, des, tadv, ma_, re_, tinf_, cfl_, dt_, aoa_deg) &
        bind(C, name="refh_setup")
    use iso_c_binding
    use flag_var
    use ns_const
    implicit none
    integer(c_int), value :: dimen, solver, itime, turb, des, tadv
    real(c_double), value :: ma_, re_, tinf_, cfl_, dt_, aoa_deg
    iflag_dimension = dimen; iflag_solver = solver; iflag_time = itime
    iflag_turbulence = turb; iflag_des = des; iflag_timeadvance = tadv
    iflag_inviscid = 4
    ma = ma_; re = re_; tinf = tinf_; cfl = cfl_; dt = dt_
    d_0 = 1.d0; u_0 = 1.d0; t_0 = 1.d0
    p_0 = d_0 * t_0 / gamma / ma / ma
    c_0 = sqrt(t_0) / ma
    cv = 1.d0 / gamma / (gamma - 1.d0) / ma / ma
    cp = 1.d0 / (gamma - 1.d0) / ma / ma
    aoa = aoa_deg * pi / 180.0d0
end subroutine

subroutine refh_scalars(out) bind(C, name="refh_scalars")
    use iso_c_binding
    use ns_const
    implicit none
    real(c_double) :: out(12)
    out = [gamma, cv, re, tinf, prl, prt, d_0, u_0, p_0, c_0, aoa, cfl]
end subroutine

subroutine refh_set_cfl(x) bind(C, name="refh_set_cfl")
    use iso_c_binding
    use ns_const
    implicit none
    real(c_double), value :: x
    cfl = x
end subroutine

subroutine refh_copy(dst, src, n)
    implicit none
    integer :: n
    real*8 :: dst(n), src(n)
    dst = src
end subroutine

!! io = 0: copy arr into the block array; io = 1: copy the block array out
subroutine refh_xfer(mm, which, arr, n, io) bind(C, name="refh_xfer")
    use iso_c_binding
    use blk_var
    implicit none
    integer(c_int), value :: mm, which, io
    integer(c_long), value :: n
    real(c_double) :: arr(n)
    select case (which)
    case (1);  call mv(blk(mm)%q)
    case (2);  call mv(blk(mm)%pri_v)
    case (3);  call mv(blk(mm)%rhsi)
    case (4);  call mv(blk(mm)%rhsv)
    case (5);  call mv(blk(mm)%rhs)
    case (6);  call mv(blk(mm)%dudx)
    case (7);  call mv(blk(mm)%tao)
    case (8);  call mv3(blk(mm)%vor)
    case (9);  call mv(blk(mm)%amu)
    case (10); call mv3(blk(mm)%dt)
    case (11); call mv(blk(mm)%dxidx)
    case (12); call mv3(blk(mm)%inv_j)
    case (13); call mv(blk(mm)%aalpha)
    case (14); call mv(blk(mm)%bbeta)
    case (15); call mv(blk(mm)%ggamma)
    case (16); call mv3(blk(mm)%dist)
    case (17); call mv3(blk(mm)%spacing)
    case (18); call mv3(blk(mm)%nut)
    case (19); call mv3(blk(mm)%sa_rhs)
    case (20); call mv5(blk(mm)%q0)
    case (21); call mv3(blk(mm)%length)
    case (22); call mv3(blk(mm)%x)
    case (23); call mv3(blk(mm)%y)
    case (24); call mv3(blk(mm)%z)
    case (25); call mv3(blk(mm)%spci)
    case (26); call mv3(blk(mm)%spcj)
    case (27); call mv3(blk(mm)%spck)
    case (28); call mv3(blk(mm)%vol)
    end select
contains
    subroutine mv(a)
        real*8, pointer :: a(:, :, :, :)
        if (size(a, kind=8) /= n) stop 'refh_xfer: size'
        if (io == 0) then
            call refh_copy(a, arr, int(n))
        else
            call refh_copy(arr, a, int(n))
        end if
    end subroutine
    subroutine mv3(a)
        real*8, pointer :: a(:, :, :)
        if (size(a, kind=8) /= n) stop 'refh_xfer: size'
        if (io == 0) then
            call refh_copy(a, arr, int(n))
        else
            call refh_copy(arr, a, int(n))
        end if
    end subroutine
    subroutine mv5(a)
        real*8, pointer :: a(:, :, :, :, :)
        if (size(a, kind=8) /= n) stop 'refh_xfer: size'
        if (io == 0) then
            call refh_copy(a, arr, int(n))
        else
            call refh_copy(arr, a, int(n))
        end if
    end subroutine
end subroutine

subroutine refh_physical_bc() bind(C, name="refh_physical_bc")
    implicit none
    call physical_bc
end subroutine

subroutine refh_update_buffer() bind(C, name="refh_update_buffer")
    implicit none
    call UpdateBuffer
end subroutine

subroutine refh_average() bind(C, name="refh_average")
    implicit none
    call average1
    call average2
end subroutine

!! init_walldist after the face coordinates of every subface are stored as
!! sub_init_readbcinfo.f90:294-400 stores them (writes result/dist_*.dat in
!! the current directory)
subroutine refh_walldist() bind(C, name="refh_walldist")
    use blk_var
    use glbindex_var
    use flag_var
    use index_var
    use mpi_var
    implicit none
    integer :: mm, s, i, j, k, a, b, fc, na, nb
    do mm = 1, blk_loop
        do s = 1, blk0(mm)%num_subface
            fc = blk0(mm)%bc0(s)%face
            if (allocated(blk0(mm)%bc0(s)%x)) then
                deallocate(blk0(mm)%bc0(s)%x, blk0(mm)%bc0(s)%y, blk0(mm)%bc0(s)%z)
            end if
            if (fc == 1 .or. fc == 2) then
                na = blk0(mm)%bc0(s)%je - blk0(mm)%bc0(s)%js + 1
                nb = blk0(mm)%bc0(s)%ke - blk0(mm)%bc0(s)%ks + 1
            else if (fc == 3 .or. fc == 4) then
                na = blk0(mm)%bc0(s)%ie - blk0(mm)%bc0(s)%is + 1
                nb = blk0(mm)%bc0(s)%ke - blk0(mm)%bc0(s)%ks + 1
            else
                na = blk0(mm)%bc0(s)%ie - blk0(mm)%bc0(s)%is + 1
                nb = blk0(mm)%bc0(s)%je - blk0(mm)%bc0(s)%js + 1
            end if
            allocate(blk0(mm)%bc0(s)%x(na, nb), blk0(mm)%bc0(s)%y(na, nb), blk0(mm)%bc0(s)%z(na, nb))
            do k = blk0(mm)%bc0(s)%ks, blk0(mm)%bc0(s)%ke
                do j = blk0(mm)%bc0(s)%js, blk0(mm)%bc0(s)%je
                    do i = blk0(mm)%bc0(s)%is, blk0(mm)%bc0(s)%ie
                        if (fc == 1 .or. fc == 2) then
                            a = j - blk0(mm)%bc0(s)%js + 1; b = k - blk0(mm)%bc0(s)%ks + 1
                        else if (fc == 3 .or. fc == 4) then
                            a = i - blk0(mm)%bc0(s)%is + 1; b = k - blk0(mm)%bc0(s)%ks + 1
                        else
                            a = i - blk0(mm)%bc0(s)%is + 1; b = j - blk0(mm)%bc0(s)%js + 1
                        end if
                        !! the face plane itself (is0/ie0, js0/je0, ks0/ke0)
                        blk0(mm)%bc0(s)%x(a, b) = blk(mm)%x(face_i(i), face_j(j), face_k(k))
                        blk0(mm)%bc0(s)%y(a, b) = blk(mm)%y(face_i(i), face_j(j), face_k(k))
                        blk0(mm)%bc0(s)%z(a, b) = blk(mm)%z(face_i(i), face_j(j), face_k(k))
                    end do
                end do
            end do
        end do
    end do
    call init_walldist
contains
    integer function face_i(i)
        integer :: i
        face_i = i
        if (fc == 1) face_i = blk(mm)%is0
        if (fc == 2) face_i = blk(mm)%ie0
    end function
    integer function face_j(j)
        integer :: j
        face_j = j
        if (fc == 3) face_j = blk(mm)%js0
        if (fc == 4) face_j = blk(mm)%je0
    end function
    integer function face_k(k)
        integer :: k
        face_k = k
        if (fc == 5) face_k = blk(mm)%ks0
        if (fc == 6) face_k = blk(mm)%ke0
    end function
end subroutine

!! init_readcontrolfile (reads input/Control.in in the current directory)
!! and the values it set, in the order of the file: integers in iv, reals
!! in rv, the three file names in sv (180 characters each)
subroutine refh_read_control(iv, rv, sv) bind(C, name="refh_read_control")
    use iso_c_binding
    use flag_var
    use ns_const
    use output_var
    implicit none
    integer(c_int) :: iv(22)
    real(c_double) :: rv(17)
    character(kind=c_char) :: sv(540)
    integer :: c
    call init_readcontrolfile
    iv = [iflag_dimension, iflag_solver, iflag_time, iflag_splittingtype, iflag_inviscid, iflag_filter, &
          iflag_timeadvance, iflag_turbulence, iflag_des, writeinterval, ntstep, iflag_init, nfile, &
          restart_timestep, nobs, iflag_acoustic, Start_Acsoutic, End_Acsoutic, iflag_avg, Start_Avg, End_Avg, &
          Avg_interval]
    rv = [cfl, dt, ma, re, tinf, aoa, scale_x, scale_y, scale_z, lref, sref, fltr, ob_dist, acoustic_length, &
          0.d0, 0.d0, 0.d0]
    do c = 1, 180
        sv(c) = grid(c:c)
        sv(180 + c) = bcinfo(c:c)
        sv(360 + c) = restartfile(c:c)
    end do
end subroutine

!! clcd_output with the reference area sref_ (it appends time, cl, cd to
!! result/clcd_history.dat in the current directory, which must exist)
subroutine refh_clcd(sref_) bind(C, name="refh_clcd")
    use iso_c_binding
    use ns_const
    implicit none
    real(c_double), value :: sref_
    sref = sref_
    timestep = 1
    call clcd_output
end subroutine

subroutine refh_update_coordinate() bind(C, name="refh_update_coordinate")
    implicit none
    call UpdateBufferCoordinate
end subroutine

!! residual writes result/resi_history.dat in the current directory
subroutine refh_residual() bind(C, name="refh_residual")
    implicit none
    call residual
end subroutine

!! One time step in the order of main.f90:343-573 (implicit1st or crank).
subroutine refh_step() bind(C, name="refh_step")
    use blk_var
    use flag_var
    use ns_const
    use index_var
    implicit none
    integer :: sub, mm
    if (iflag_timeadvance == iflag_1stimplicit) then
        do mm = 1, blk_loop
            call stage(mm)
            call implicit1st(blk(mm)%q, blk(mm)%dt, blk(mm)%pri_v, blk(mm)%rhs, blk(mm)%rhsi, blk(mm)%rhsv, &
                             blk(mm)%dxidx, blk(mm)%inv_j, blk(mm)%is, blk(mm)%ie, blk(mm)%js, blk(mm)%je, &
                             blk(mm)%ks, blk(mm)%ke, blk(mm)%is0, blk(mm)%ie0, blk(mm)%js0, blk(mm)%je0, &
                             blk(mm)%ks0, blk(mm)%ke0, blk(mm)%is1, blk(mm)%ie1, blk(mm)%js1, blk(mm)%je1, &
                             blk(mm)%ks1, blk(mm)%ke1, gamma, cfl, blk(mm)%x, blk(mm)%y, blk(mm)%z)
            if (iflag_turbulence == iflag_sa .and. iflag_solver == iflag_nssolver) call samodel(mm)
        end do
        call refresh()
        do mm = 1, blk_loop
            call get_primitivevariables(mm, blk(mm)%pri_v, blk(mm)%q, gamma, cv, blk(mm)%is, blk(mm)%ie, &
                                        blk(mm)%js, blk(mm)%je, blk(mm)%ks, blk(mm)%ke)
        end do
    else if (iflag_timeadvance == iflag_2ndcrank) then
        do sub = 1, 3
            call refresh()
            do mm = 1, blk_loop
                call stage(mm)
                call crank2nd(blk(mm)%q, blk(mm)%q0, blk(mm)%dt, blk(mm)%pri_v, blk(mm)%rhs, blk(mm)%rhsi, &
                              blk(mm)%rhsv, blk(mm)%dxidx, blk(mm)%inv_j, blk(mm)%is, blk(mm)%ie, blk(mm)%js, &
                              blk(mm)%je, blk(mm)%ks, blk(mm)%ke, blk(mm)%is0, blk(mm)%ie0, blk(mm)%js0, &
                              blk(mm)%je0, blk(mm)%ks0, blk(mm)%ke0, blk(mm)%is1, blk(mm)%ie1, blk(mm)%js1, &
                              blk(mm)%je1, blk(mm)%ks1, blk(mm)%ke1, gamma, mm, sub)
            end do
        end do
        do mm = 1, blk_loop
            call shift_q0(mm)
        end do
        do mm = 1, blk_loop
            if (iflag_turbulence == iflag_sa .and. iflag_solver == iflag_nssolver) call samodel(mm)
        end do
        call refresh()
        do mm = 1, blk_loop
            call get_primitivevariables(mm, blk(mm)%pri_v, blk(mm)%q, gamma, cv, blk(mm)%is, blk(mm)%ie, &
                                        blk(mm)%js, blk(mm)%je, blk(mm)%ks, blk(mm)%ke)
        end do
    end if
contains
    subroutine refresh()
        call physical_bc
        call UpdateBuffer
        call average1
        call average2
    end subroutine
    subroutine stage(m1)
        integer :: m1
        call get_primitivevariables(m1, blk(m1)%pri_v, blk(m1)%q, gamma, cv, blk(m1)%is, blk(m1)%ie, &
                                    blk(m1)%js, blk(m1)%je, blk(m1)%ks, blk(m1)%ke)
        call inviscid_flux_1(blk(m1)%rhsi, blk(m1)%q, blk(m1)%dxidx, blk(m1)%inv_j, &
                             blk(m1)%is, blk(m1)%ie, blk(m1)%js, blk(m1)%je, blk(m1)%ks, blk(m1)%ke, &
                             blk(m1)%is0, blk(m1)%ie0, blk(m1)%js0, blk(m1)%je0, blk(m1)%ks0, blk(m1)%ke0, &
                             blk(m1)%is1, blk(m1)%ie1, blk(m1)%js1, blk(m1)%je1, blk(m1)%ks1, blk(m1)%ke1, gamma)
        if (iflag_solver == iflag_nssolver) then
            call viscous_flux(blk(m1)%rhsv, blk(m1)%pri_v, blk(m1)%amu, blk(m1)%vor, blk(m1)%dudx, blk(m1)%tao, &
                              blk(m1)%dxidx, blk(m1)%inv_j, tinf, prl, prt, gamma, cv, re, &
                              blk(m1)%is, blk(m1)%ie, blk(m1)%js, blk(m1)%je, blk(m1)%ks, blk(m1)%ke, &
                              blk(m1)%is0, blk(m1)%ie0, blk(m1)%js0, blk(m1)%je0, blk(m1)%ks0, blk(m1)%ke0, &
                              blk(m1)%is1, blk(m1)%ie1, blk(m1)%js1, blk(m1)%je1, blk(m1)%ks1, blk(m1)%ke1, m1)
        end if
        call find_dt(blk(m1)%dt, dtmax, dtmin, cflmax, cflmin, blk(m1)%pri_v, blk(m1)%amu, blk(m1)%dxidx, &
                     re, cfl, blk(m1)%is, blk(m1)%ie, blk(m1)%js, blk(m1)%je, blk(m1)%ks, blk(m1)%ke, &
                     iflag_time, iflag_steady, iflag_unsteady)
    end subroutine
    subroutine samodel(m1)
        integer :: m1
        call sa_model(blk(m1)%nut, blk(m1)%sa_rhs, blk(m1)%dt, blk(m1)%pri_v, blk(m1)%amu, blk(m1)%dudx, &
                      blk(m1)%vor, blk(m1)%dxidx, blk(m1)%inv_j, blk(m1)%aalpha, blk(m1)%bbeta, blk(m1)%ggamma, &
                      blk(m1)%length, blk(m1)%dist, blk(m1)%spacing, re, cfl, &
                      blk(m1)%is, blk(m1)%ie, blk(m1)%js, blk(m1)%je, blk(m1)%ks, blk(m1)%ke, &
                      blk(m1)%is0, blk(m1)%ie0, blk(m1)%js0, blk(m1)%je0, blk(m1)%ks0, blk(m1)%ke0, &
                      blk(m1)%is1, blk(m1)%ie1, blk(m1)%js1, blk(m1)%je1, blk(m1)%ks1, blk(m1)%ke1)
    end subroutine
    !! the q0 history shift of main.f90:692-711 (update_q0 lives in the
    !! program file, which is not linked into this library)
    subroutine shift_q0(m1)
        integer :: m1, i, j, k
        do k = blk(m1)%ks0, blk(m1)%ke0
        do j = blk(m1)%js0, blk(m1)%je0
        do i = blk(m1)%is0, blk(m1)%ie0
            blk(m1)%q0(-1, :, i, j, k) = blk(m1)%q0(0, :, i, j, k)
            blk(m1)%q0(0, :, i, j, k) = blk(m1)%q(:, i, j, k)
        end do
        end do
        end do
    end subroutine
end subroutine

subroutine refh_stats(out) bind(C, name="refh_stats")
    use iso_c_binding
    use ns_const
    implicit none
    real(c_double) :: out(4)
    out = [dtmax, dtmin, cflmax, cflmin]
end subroutine
