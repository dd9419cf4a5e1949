!! A Fortran host of the residual-and-update path on the GPU.
!!
!! Reads a case from the reference's own files in the current directory --
!! input/Control.in (sub_init_readcontrolfile.f90:27-47), the grid file
!! (sub_init_allocatememory.f90:36-71,388-417), the bcinfo file
!! (sub_init_readbcinfo.f90:45-104) and, for iflag_init = 1, the restart file
!! result/flowfield_<restart_timestep>_0000.dat (sub_init_flowfield.f90:
!! 126-262) -- and runs ntstep steps through libaerocfd_amd.so, calling it
!! only through the ISO_C_BINDING module of include/aerocfd_amd.f90.  It
!! stands in for main.f90 on one process: the coordinate halo exchange and
!! the grid pre-processing run on the device (acfd_exchange_coordinates,
!! acfd_grid_metrics), the time loop is acfd_step, and it writes the
!! reference's outputs, result/resi_history.dat every step
!! (sub_output_residual.f90:62-64) and result/flowfield_<step>_0000.dat every
!! writeinterval steps (sub_output_flowfield.f90:30-81).
!!
!! Usage: acfd_host [lusgs_kernel batch_blocks use_graph precision]
!! (defaults 2 1 1 0: batched tile sweeps and stages, graph replay, FP64).
program acfd_host
  use iso_c_binding
  use aerocfd_amd
  implicit none

  type block_t
     integer :: r(18)
     real(c_double), allocatable :: xyz(:, :, :, :)   ! x, y, z over the extended extents
     real(c_double), allocatable :: q(:, :, :, :), amu(:, :, :, :), pri(:, :, :, :)
     real(c_double), allocatable :: dt(:, :, :), nut(:, :, :)
  end type block_t

  integer :: idim, isolver, itime, isplit, iinv, ifilter, itadv, iturb, ides
  integer :: writeinterval, ntstep, iinit, nfile, rstep
  real(c_double) :: cfl, dt, ma, re, tinf, aoa, sx, sy, sz, lref, sref, fltr
  character(len=180) :: gridfile, bcfile, restartfile
  character(len=32) :: arg

  integer :: nblock, m, ns, step, i, j, k
  integer, allocatable :: dims(:, :), nsub(:)
  type(acfd_subface), allocatable :: subs(:)
  type(block_t), allocatable, target :: blk(:)
  type(acfd_config) :: cfg
  type(c_ptr) :: ctx
  type(c_ptr), allocatable, target :: xyz_ptr(:)
  integer :: opt(4)
  real(c_double) :: res(7), gamma, cv, d0, u0, t0, p0, tt, nutinf, aoa_rad
  real(c_double), parameter :: pi = 3.1415926535897932d0
  character(len=64) :: fname

  opt = [2, 1, 1, 0]
  do i = 1, min(command_argument_count(), 4)
     call get_command_argument(i, arg)
     read(arg, *) opt(i)
  end do

  ! ------------------------------------------------------------ Control.in
  open(10, file="input/Control.in", status="old", action="read")
  call skip2(10)
  read(10, *) idim, isolver, itime, cfl, dt
  call skip2(10)
  read(10, *) ma, re, tinf, aoa
  call skip2(10)
  read(10, *) sx, sy, sz, lref, sref, fltr
  call skip2(10)
  read(10, *) isplit, iinv, ifilter, itadv, iturb, ides
  call skip2(10)
  read(10, *) writeinterval, ntstep, gridfile, bcfile, iinit, restartfile, nfile, rstep
  close(10)
  if (idim /= 3) call die("only 3-D cases are supported")

  ! ------------------------------------------------------------------ grid
  open(11, file=trim(gridfile), status="old", action="read")
  read(11, *) nblock
  allocate(dims(3, nblock), nsub(nblock), blk(nblock), xyz_ptr(nblock))
  do m = 1, nblock
     read(11, *) dims(:, m)
  end do

  ! ---------------------------------------------------------------- bcinfo
  open(12, file=trim(bcfile), status="old", action="read")
  read(12, *)
  read(12, *)
  read(12, *)
  allocate(subs(0))
  do m = 1, nblock
     read(12, *)
     read(12, *)
     read(12, *) ns
     read(12, *)
     nsub(m) = ns
     call add_subfaces(12, ns)
     read(12, *)
  end do
  close(12)

  ! --------------------------------------------------------------- context
  cfg%ma = ma; cfg%re = re; cfg%tinf = tinf; cfg%aoa_deg = aoa; cfg%cfl = cfl; cfg%dt = dt
  cfg%prl = 0.72d0; cfg%prt = 0.90d0
  cfg%iflag_solver = isolver; cfg%iflag_time = itime; cfg%iflag_turbulence = iturb; cfg%iflag_des = ides
  cfg%iflag_timeadvance = itadv; cfg%precision = opt(4); cfg%lusgs_kernel = opt(1)
  cfg%inviscid_scheme = 0; cfg%use_graph = opt(3); cfg%batch_blocks = opt(2); cfg%reserved = 0
  call check(acfd_create(cfg, ctx), "acfd_create")
  call check(acfd_set_topology(ctx, nblock, dims, nsub, subs), "acfd_set_topology")

  ! coordinates of the physical range, scaled as the reference scales them
  do m = 1, nblock
     call check(acfd_block_ranges(ctx, m, blk(m)%r), "acfd_block_ranges")
     associate (r => blk(m)%r)
       allocate(blk(m)%xyz(r(1):r(2), r(3):r(4), r(5):r(6), 3))
       blk(m)%xyz = 0.d0
       read(11, *) (((blk(m)%xyz(i, j, k, 1), i = r(7), r(8)), j = r(9), r(10)), k = r(11), r(12)), &
                   (((blk(m)%xyz(i, j, k, 2), i = r(7), r(8)), j = r(9), r(10)), k = r(11), r(12)), &
                   (((blk(m)%xyz(i, j, k, 3), i = r(7), r(8)), j = r(9), r(10)), k = r(11), r(12))
       blk(m)%xyz(:, :, :, 1) = blk(m)%xyz(:, :, :, 1) * sx / lref
       blk(m)%xyz(:, :, :, 2) = blk(m)%xyz(:, :, :, 2) * sy / lref
       blk(m)%xyz(:, :, :, 3) = blk(m)%xyz(:, :, :, 3) * sz / lref
     end associate
     xyz_ptr(m) = c_loc(blk(m)%xyz)
  end do
  close(11)
  ! halo coordinates from the neighbours, then metrics, edge lengths and
  ! (SA) wall distance on the device
  call check(acfd_exchange_coordinates(ctx, xyz_ptr, c_null_ptr), "acfd_exchange_coordinates")
  call check(acfd_grid_metrics(ctx, xyz_ptr), "acfd_grid_metrics")

  ! ------------------------------------------------- initial flow field
  gamma = 1.4d0
  d0 = 1.d0; u0 = 1.d0; t0 = 1.d0
  p0 = d0 * t0 / gamma / ma / ma
  cv = 1.d0 / gamma / (gamma - 1.d0) / ma / ma
  aoa_rad = aoa * pi / 180.0d0
  nutinf = 1.341946d0
  if (iinit == 1) then
     write(fname, "('result/flowfield_',I10.10,'_',I4.4,'.dat')") rstep, 0
     open(13, file=trim(fname), status="old", action="read")
  end if
  do m = 1, nblock
     associate (r => blk(m)%r)
       allocate(blk(m)%q(5, r(1):r(2), r(3):r(4), r(5):r(6)), blk(m)%amu(3, r(1):r(2), r(3):r(4), r(5):r(6)))
       allocate(blk(m)%pri(7, r(1):r(2), r(3):r(4), r(5):r(6)))
       allocate(blk(m)%dt(r(1):r(2), r(3):r(4), r(5):r(6)), blk(m)%nut(r(1):r(2), r(3):r(4), r(5):r(6)))
       blk(m)%q = 0.d0; blk(m)%amu = 0.d0; blk(m)%dt = 0.d0; blk(m)%nut = 0.d0
       if (iinit == 0) then
          ! freestream on the extended range (sub_init_flowfield.f90:43-110)
          blk(m)%q(1, :, :, :) = d0
          blk(m)%q(2, :, :, :) = d0 * u0 * cos(aoa_rad)
          blk(m)%q(3, :, :, :) = d0 * u0 * sin(aoa_rad)
          blk(m)%q(5, :, :, :) = d0 * t0 * cv + 0.5d0 * d0 * u0 * u0
          blk(m)%dt = dt
          blk(m)%amu(1, :, :, :) = 1.d0
          if (iturb == 2) then
             blk(m)%nut = nutinf
             blk(m)%amu(2, :, :, :) = 0.009d0
          end if
          blk(m)%amu(3, :, :, :) = blk(m)%amu(1, :, :, :) + blk(m)%amu(2, :, :, :)
       else
          call read_restart_block(13, blk(m))
       end if
     end associate
     call check(acfd_upload(ctx, m, ACFD_Q, blk(m)%q), "upload q")
     call check(acfd_upload(ctx, m, ACFD_AMU, blk(m)%amu), "upload amu")
     call check(acfd_upload(ctx, m, ACFD_DT, blk(m)%dt), "upload dt")
     if (iturb == 2) call check(acfd_upload(ctx, m, ACFD_NUT, blk(m)%nut), "upload nut")
  end do
  if (iinit == 1) close(13)

  ! boundary conditions and primitive variables, then the pre-loop refresh
  ! (main.f90:249-257, 309-318)
  call check(acfd_run_stage(ctx, ACFD_STAGE_PHYSICAL_BC), "physical_bc")
  call check(acfd_run_stage(ctx, ACFD_STAGE_GETPRI), "getpri")
  call check(acfd_prepare(ctx), "acfd_prepare")

  open(14, file="result/resi_history.dat", status="replace", action="write")
  write(14, *) 'variables="timestep","rou","momentum u","momentum v","momentum w","energy","tur1","tur2"'
  do step = 1, ntstep
     call check(acfd_step(ctx, 1), "acfd_step")
     call check(acfd_residual(ctx, res), "acfd_residual")
     write(14, "(I6.6,6E11.4)") step, res(1), res(2), res(3), res(4), res(5), res(6)
     if (mod(step, writeinterval) == 0) call write_flowfield(step)
  end do
  close(14)
  call check(acfd_destroy(ctx), "acfd_destroy")
  write(*, "(A,I0,A,I0,A)") "acfd_host: ", ntstep, " steps of ", nblock, " blocks done"

contains

  subroutine skip2(u)
    integer, intent(in) :: u
    read(u, *)
    read(u, *)
  end subroutine skip2

  subroutine add_subfaces(u, n)
    integer, intent(in) :: u, n
    type(acfd_subface), allocatable :: grown(:)
    integer :: s, fwh, rec(11)
    allocate(grown(size(subs) + n))
    grown(1:size(subs)) = subs
    do s = 1, n
       read(u, *) rec, fwh
       grown(size(subs) + s) = acfd_subface(rec(1), rec(2), rec(3), rec(4), rec(5), rec(6), rec(7), rec(8), &
                                            rec(9), rec(10), rec(11))
    end do
    call move_alloc(grown, subs)
  end subroutine add_subfaces

  !! One block of a restart file: x y z d u v w p nut per node of the
  !! physical range, turned into q as sub_init_flowfield.f90:206-258 does.
  subroutine read_restart_block(u, b)
    integer, intent(in) :: u
    type(block_t), intent(inout) :: b
    real(c_double) :: x, y, z, d, uu, vv, ww, p, tur, t
    integer :: i, j, k
    read(u, *)
    read(u, *)
    do k = b%r(11), b%r(12)
       do j = b%r(9), b%r(10)
          do i = b%r(7), b%r(8)
             read(u, *) x, y, z, d, uu, vv, ww, p, tur
             t = p / d * gamma * ma * ma
             b%q(1, i, j, k) = d
             b%q(2, i, j, k) = d * uu
             b%q(3, i, j, k) = d * vv
             b%q(4, i, j, k) = d * ww
             b%q(5, i, j, k) = d * t * cv + 0.5d0 * d * (uu * uu + vv * vv + ww * ww)
             b%dt(i, j, k) = dt
             b%amu(1, i, j, k) = 1.d0
             if (iturb == 2) b%nut(i, j, k) = tur
          end do
       end do
    end do
  end subroutine read_restart_block

  !! The reference's flowfield file of one step (one process: rank 0).
  subroutine write_flowfield(step)
    integer, intent(in) :: step
    character(len=64) :: fn
    integer :: m, i, j, k
    write(fn, "('result/flowfield_',I10.10,'_',I4.4,'.dat')") step, 0
    open(15, file=trim(fn), status="replace", action="write")
    do m = 1, nblock
       call check(acfd_download(ctx, m, ACFD_PRI_V, blk(m)%pri), "download pri_v")
       if (iturb == 2) call check(acfd_download(ctx, m, ACFD_NUT, blk(m)%nut), "download nut")
       associate (r => blk(m)%r)
         write(15, *) "variables = x,y,z,d,u,v,w,p,nut"
         write(15, *) "zone i= ", r(8) - r(7) + 1, "j= ", r(10) - r(9) + 1, "k= ", r(12) - r(11) + 1, &
                      "datapacking=point"
         do k = r(11), r(12)
            do j = r(9), r(10)
               do i = r(7), r(8)
                  write(15, *) blk(m)%xyz(i, j, k, 1), blk(m)%xyz(i, j, k, 2), blk(m)%xyz(i, j, k, 3), &
                               blk(m)%pri(1:5, i, j, k), blk(m)%nut(i, j, k)
               end do
            end do
         end do
       end associate
    end do
    close(15)
  end subroutine write_flowfield

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= ACFD_OK) then
       write(*, "(A,A,A,I0,A,A)") "acfd_host: ", what, " failed (", rc, "): ", acfd_error_message()
       stop 1
    end if
  end subroutine check

  subroutine die(msg)
    character(len=*), intent(in) :: msg
    write(*, "(A,A)") "acfd_host: ", msg
    stop 2
  end subroutine die

end program acfd_host
