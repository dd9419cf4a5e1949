!! ISO_C_BINDING interface to libaerocfd_amd.so (the C ABI of aerocfd_amd.h).
!!
!! The derived types mirror acfd_config and acfd_subface field by field; the
!! functions are the entry points a Fortran host of the reference calls in
!! place of its time-loop body (main.f90:342-613), its start-up grid
!! pre-processing (main.f90:163-209) and its force output
!! (sub_output_forceintegration.f90:8).  Every array argument is passed as
!! the reference stores it (column-major, component first, extended
!! extents), e.g. blk(m)%q(5,is:ie,js:je,ks:ke).
module aerocfd_amd
  use iso_c_binding
  implicit none

  type, bind(C) :: acfd_config
     real(c_double) :: ma, re, tinf, aoa_deg, cfl, dt, prl, prt
     integer(c_int) :: iflag_solver, iflag_time, iflag_turbulence, iflag_des
     integer(c_int) :: iflag_timeadvance, precision, lusgs_kernel
     integer(c_int) :: inviscid_scheme, use_graph, batch_blocks
     integer(c_int) :: reserved(2)
  end type acfd_config

  type, bind(C) :: acfd_subface
     integer(c_int) :: f_no, face, is, ie, js, je, ks, ke, blk_t, subface_t, ori
  end type acfd_subface

  integer(c_int), parameter :: ACFD_OK = 0, ACFD_EINVAL = -1, ACFD_EDEVICE = -2, ACFD_ESTATE = -3, &
       ACFD_ENUMERIC = -4

  integer(c_int), parameter :: ACFD_Q = 1, ACFD_PRI_V = 2, ACFD_RHSI = 3, ACFD_RHSV = 4, ACFD_RHS = 5, &
       ACFD_DUDX = 6, ACFD_TAO = 7, ACFD_VOR = 8, ACFD_AMU = 9, ACFD_DT = 10, ACFD_DXIDX = 11, &
       ACFD_INV_J = 12, ACFD_AALPHA = 13, ACFD_BBETA = 14, ACFD_GGAMMA = 15, ACFD_DIST = 16, &
       ACFD_SPACING = 17, ACFD_NUT = 18, ACFD_SA_RHS = 19, ACFD_Q0 = 20, ACFD_LENGTH = 21, &
       ACFD_X = 22, ACFD_Y = 23, ACFD_Z = 24, ACFD_SPCI = 25, ACFD_SPCJ = 26, ACFD_SPCK = 27

  integer(c_int), parameter :: ACFD_STAGE_GETPRI = 1, ACFD_STAGE_INVISCID = 2, ACFD_STAGE_VISCOUS = 3, &
       ACFD_STAGE_FIND_DT = 4, ACFD_STAGE_IMPLICIT1ST = 5, ACFD_STAGE_SA = 6, ACFD_STAGE_PHYSICAL_BC = 7, &
       ACFD_STAGE_UPDATE_BUFFER = 8, ACFD_STAGE_AVERAGE = 9, ACFD_STAGE_CRANK2ND = 10, &
       ACFD_STAGE_UPDATE_Q0 = 11

  interface
     type(c_ptr) function acfd_last_error() bind(C, name="acfd_last_error")
       import
     end function acfd_last_error
     integer(c_int) function acfd_create(cfg, ctx) bind(C, name="acfd_create")
       import
       type(acfd_config), intent(in) :: cfg
       type(c_ptr) :: ctx
     end function acfd_create
     integer(c_int) function acfd_destroy(ctx) bind(C, name="acfd_destroy")
       import
       type(c_ptr), value :: ctx
     end function acfd_destroy
     integer(c_int) function acfd_set_topology(ctx, nblk, dims, nsub, subs) bind(C, name="acfd_set_topology")
       import
       type(c_ptr), value :: ctx
       integer(c_int), value :: nblk
       integer(c_int), intent(in) :: dims(*), nsub(*)
       type(acfd_subface), intent(in) :: subs(*)
     end function acfd_set_topology
     integer(c_int) function acfd_block_ranges(ctx, m, ranges) bind(C, name="acfd_block_ranges")
       import
       type(c_ptr), value :: ctx
       integer(c_int), value :: m
       integer(c_int) :: ranges(18)
     end function acfd_block_ranges
     integer(c_int) function acfd_upload(ctx, m, field, host) bind(C, name="acfd_upload")
       import
       type(c_ptr), value :: ctx
       integer(c_int), value :: m, field
       real(c_double), intent(in) :: host(*)
     end function acfd_upload
     integer(c_int) function acfd_download(ctx, m, field, host) bind(C, name="acfd_download")
       import
       type(c_ptr), value :: ctx
       integer(c_int), value :: m, field
       real(c_double) :: host(*)
     end function acfd_download
     integer(c_int) function acfd_run_stage(ctx, stage) bind(C, name="acfd_run_stage")
       import
       type(c_ptr), value :: ctx
       integer(c_int), value :: stage
     end function acfd_run_stage
     integer(c_int) function acfd_prepare(ctx) bind(C, name="acfd_prepare")
       import
       type(c_ptr), value :: ctx
     end function acfd_prepare
     integer(c_int) function acfd_step(ctx, nsteps) bind(C, name="acfd_step")
       import
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
     end function acfd_step
     integer(c_int) function acfd_residual(ctx, out) bind(C, name="acfd_residual")
       import
       type(c_ptr), value :: ctx
       real(c_double) :: out(7)
     end function acfd_residual
     integer(c_int) function acfd_stats(ctx, out) bind(C, name="acfd_stats")
       import
       type(c_ptr), value :: ctx
       real(c_double) :: out(5)
     end function acfd_stats
     integer(c_int) function acfd_sync(ctx) bind(C, name="acfd_sync")
       import
       type(c_ptr), value :: ctx
     end function acfd_sync
     ! xyz(m) = c_loc of block m's x, y, z stored one after the other over the
     ! extended extents; shift = c_null_ptr or one translation per subface
     integer(c_int) function acfd_exchange_coordinates(ctx, xyz, shift) bind(C, name="acfd_exchange_coordinates")
       import
       type(c_ptr), value :: ctx
       type(c_ptr), intent(in) :: xyz(*)
       type(c_ptr), value :: shift
     end function acfd_exchange_coordinates
     integer(c_int) function acfd_grid_metrics(ctx, xyz) bind(C, name="acfd_grid_metrics")
       import
       type(c_ptr), value :: ctx
       type(c_ptr), intent(in) :: xyz(*)
     end function acfd_grid_metrics
     integer(c_int) function acfd_forces(ctx, sref, out) bind(C, name="acfd_forces")
       import
       type(c_ptr), value :: ctx
       real(c_double), value :: sref
       real(c_double) :: out(8)
     end function acfd_forces
     ! stateless routines (one block, its 18 index ranges)
     integer(c_int) function acfd_getpri(ranges, gamma, cv, q, pri_v) bind(C, name="acfd_getpri")
       import
       integer(c_int), intent(in) :: ranges(18)
       real(c_double), value :: gamma, cv
       real(c_double), intent(in) :: q(*)
       real(c_double) :: pri_v(*)
     end function acfd_getpri
     integer(c_int) function acfd_inviscid_rhs(cfg, ranges, q, dxidx, inv_j, rhsi) bind(C, name="acfd_inviscid_rhs")
       import
       type(acfd_config), intent(in) :: cfg
       integer(c_int), intent(in) :: ranges(18)
       real(c_double), intent(in) :: q(*), dxidx(*), inv_j(*)
       real(c_double) :: rhsi(*)
     end function acfd_inviscid_rhs
     integer(c_int) function acfd_lusgs(cfg, ranges, q, dt, pri_v, rhsi, rhsv, dxidx, inv_j, cfl, rhs) &
          bind(C, name="acfd_lusgs")
       import
       type(acfd_config), intent(in) :: cfg
       integer(c_int), intent(in) :: ranges(18)
       real(c_double) :: q(*), rhs(*)
       real(c_double), intent(in) :: dt(*), pri_v(*), rhsi(*), rhsv(*), dxidx(*), inv_j(*)
       real(c_double), value :: cfl
     end function acfd_lusgs
  end interface

contains

  !! The message of the last error, as a Fortran string.
  function acfd_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: n
    cp = acfd_last_error()
    if (.not. c_associated(cp)) then
       msg = ""
       return
    end if
    call c_f_pointer(cp, p, [4096])
    n = 0
    do while (n < 4096)
       if (p(n + 1) == c_null_char) exit
       n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function acfd_error_message

end module aerocfd_amd
