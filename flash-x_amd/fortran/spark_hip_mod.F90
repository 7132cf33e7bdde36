!! spark_hip_mod.F90 -- ISO_C_BINDING interface of libspark_hip.so, the
!! MI355X-native Spark unsplit-hydro RK stage (C declarations in
!! include/spark_hip.h).
!!
!! The interfaces bind the C entry points one to one; spark_check maps a
!! non-zero status to an abort with the library's message, as the
!! reference reports failures through Driver_abort
!! (Flash-X/source/Driver/Driver_abort.F90).  Inside Flash-X the abort hook
!! is pointed at Driver_abort (see Hydro_sparkHip.F90); standalone it is
!! ERROR STOP.
module spark_hip
  use iso_c_binding
  implicit none
  private

  integer, parameter, public :: SPARK_OK = 0
  integer, parameter, public :: SPARK_NPROP = 11
  integer, parameter, public :: SPARK_MAX_SCALARS = 8
  integer, parameter, public :: SPARK_RECON_WENO5 = 0, SPARK_RECON_MP5 = 1, SPARK_RECON_PLM = 2, &
                                SPARK_RECON_FOG = 3
  integer, parameter, public :: SPARK_RIEMANN_HLLC = 0, SPARK_RIEMANN_HLL = 1
  integer, parameter, public :: SPARK_FORM_PLAIN_FP64 = 0, SPARK_FORM_FP32_FLUX = 1, &
                                SPARK_FORM_BATCHED_FP64 = 2
  integer, parameter, public :: SPARK_GUARD_TELESCOPE = 0, SPARK_GUARD_REFILL = 1
  integer, parameter, public :: SPARK_OPT_GRAPH = 1, SPARK_OPT_AUTOTUNE = 2, SPARK_OPT_SAME_LEVEL_HALOS = 3
  integer, parameter, public :: SPARK_BC_OUTFLOW = 0, SPARK_BC_PERIODIC = 1, SPARK_BC_REFLECT = 2
  integer, parameter, public :: SPARK_EOS_GAMMA = 0, SPARK_EOS_MULTIGAMMA = 1
  integer, parameter, public :: SPARK_GEOM_CARTESIAN = 0, SPARK_GEOM_CYLINDRICAL = 1, SPARK_GEOM_SPHERICAL = 2

  !! struct spark_params (same member order and types)
  type, bind(C), public :: spark_params_t
     real(c_double) :: gamma, cfl, cvisc, smlrho, smallp, smalle, smallx, eintSwitch, tiny, &
                       gasConst, abar
     real(c_double) :: grav(3)
     integer(c_int) :: ndim, nguard, nstencil, rk_order, recon, riemann
     integer(c_int) :: use_flattening, use_hybrid_riemann, nscalars, use_gravity, flux_correct
     integer(c_int) :: form, guard_mode, nunk_vars
     integer(c_int) :: var_index(SPARK_NPROP + SPARK_MAX_SCALARS)
     integer(c_int) :: nspecies, eos, reserved
     real(c_double) :: species_a(SPARK_MAX_SCALARS), species_gamma(SPARK_MAX_SCALARS)
  end type spark_params_t

  !! struct spark_kernel_stat (per-kernel HIP-event timing)
  type, bind(C), public :: spark_kernel_stat_t
     character(kind=c_char) :: name(32)
     integer(c_long_long) :: launches
     real(c_double) :: total_ms, bytes, flops
  end type spark_kernel_stat_t

  !! struct spark_block_desc
  type, bind(C), public :: spark_block_desc_t
     integer(c_int) :: nx, ny, nz, level
     real(c_double) :: dx, dy, dz
  end type spark_block_desc_t

  public :: spark_params_default, spark_init, spark_finalize, spark_batch_create, spark_batch_destroy, &
            spark_pack, spark_unpack, spark_set_uniform_grid, spark_fill_guardcells, spark_compute_dt, &
            spark_hydro_step, spark_step_begin, spark_stage, spark_step_end, spark_advance, &
            spark_get_flux_registers, spark_set_flux_corrections, spark_flux_correct, &
            spark_set_option, spark_batched_tiles, spark_synchronize, spark_get_counters, &
            spark_last_error, spark_profile, spark_profile_read, spark_set_amr_grid, spark_flux_exchange, &
            spark_set_dt_control, spark_set_geometry, spark_pack_noalloc
  public :: spark_check, spark_set_abort_hook, spark_error_message

  interface
     integer(c_int) function spark_params_default(p) bind(C, name="spark_params_default")
       import :: c_int, spark_params_t
       type(spark_params_t), intent(out) :: p
     end function spark_params_default

     integer(c_int) function spark_init(device) bind(C, name="spark_init")
       import :: c_int
       integer(c_int), value :: device
     end function spark_init

     integer(c_int) function spark_finalize() bind(C, name="spark_finalize")
       import :: c_int
     end function spark_finalize

     integer(c_int) function spark_batch_create(p, nblocks, blocks, batch) bind(C, name="spark_batch_create")
       import :: c_int, c_ptr, spark_params_t, spark_block_desc_t
       type(spark_params_t), intent(in) :: p
       integer(c_int), value :: nblocks
       type(spark_block_desc_t), intent(in) :: blocks(*)
       type(c_ptr), intent(out) :: batch
     end function spark_batch_create

     integer(c_int) function spark_batch_destroy(batch) bind(C, name="spark_batch_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
     end function spark_batch_destroy

     !! unk_blocks(i) = c_loc(unk(1,1,1,1,blockID)) of the i-th block
     integer(c_int) function spark_pack(batch, unk_blocks) bind(C, name="spark_pack")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
       type(c_ptr), intent(in) :: unk_blocks(*)
     end function spark_pack

     !! spark_pack through a fixed device staging area (no device allocation)
     integer(c_int) function spark_pack_noalloc(batch, unk_blocks) bind(C, name="spark_pack_noalloc")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
       type(c_ptr), intent(in) :: unk_blocks(*)
     end function spark_pack_noalloc

     integer(c_int) function spark_unpack(batch, unk_blocks, interior_only, which) bind(C, name="spark_unpack")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
       type(c_ptr), intent(in) :: unk_blocks(*)
       integer(c_int), value :: interior_only, which
     end function spark_unpack

     integer(c_int) function spark_set_uniform_grid(batch, nbx, nby, nbz) bind(C, name="spark_set_uniform_grid")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
       integer(c_int), value :: nbx, nby, nbz
     end function spark_set_uniform_grid

     !! origins(3*nblocks): first interior cell of each block, global index at its level;
     !! bc(3): SPARK_BC_OUTFLOW, SPARK_BC_PERIODIC or SPARK_BC_REFLECT per dimension
     integer(c_int) function spark_set_amr_grid(batch, nbx0, nby0, nbz0, origins, bc) &
          bind(C, name="spark_set_amr_grid")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
       integer(c_int), value :: nbx0, nby0, nbz0
       integer(c_int), intent(in) :: origins(*)
       integer(c_int), intent(in) :: bc(3)
     end function spark_set_amr_grid

     integer(c_int) function spark_flux_exchange(batch) bind(C, name="spark_flux_exchange")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
     end function spark_flux_exchange

     integer(c_int) function spark_fill_guardcells(batch) bind(C, name="spark_fill_guardcells")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
     end function spark_fill_guardcells

     integer(c_int) function spark_compute_dt(batch, dt, minloc) bind(C, name="spark_compute_dt")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: batch
       real(c_double), intent(out) :: dt
       integer(c_int), intent(out) :: minloc(5)
     end function spark_compute_dt

     integer(c_int) function spark_hydro_step(batch, dt, dtOld) bind(C, name="spark_hydro_step")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: batch
       real(c_double), value :: dt, dtOld
     end function spark_hydro_step

     integer(c_int) function spark_step_begin(batch) bind(C, name="spark_step_begin")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
     end function spark_step_begin

     integer(c_int) function spark_stage(batch, stage, dt, ngoff) bind(C, name="spark_stage")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: batch
       integer(c_int), value :: stage
       real(c_double), value :: dt
       integer(c_int), value :: ngoff
     end function spark_stage

     integer(c_int) function spark_step_end(batch) bind(C, name="spark_step_end")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
     end function spark_step_end

     integer(c_int) function spark_advance(batch, nsteps, dts) bind(C, name="spark_advance")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: batch
       integer(c_int), value :: nsteps
       real(c_double), intent(out) :: dts(*)
     end function spark_advance

     integer(c_int) function spark_set_geometry(batch, geometry, r_center, x_face_area, cell_volume, r_left, &
          r_right) bind(C, name="spark_set_geometry")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
       integer(c_int), value :: geometry
       type(c_ptr), intent(in) :: r_center(*), x_face_area(*), cell_volume(*), r_left(*), r_right(*)
     end function spark_set_geometry

     integer(c_int) function spark_set_dt_control(batch, dt_init, change_factor, dt_min, dt_max) &
          bind(C, name="spark_set_dt_control")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: batch
       real(c_double), value :: dt_init, change_factor, dt_min, dt_max
     end function spark_set_dt_control

     integer(c_int) function spark_get_flux_registers(batch, fbx, fby, fbz) bind(C, name="spark_get_flux_registers")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
       type(c_ptr), intent(in) :: fbx(*), fby(*), fbz(*)
     end function spark_get_flux_registers

     integer(c_int) function spark_set_flux_corrections(batch, fbx, fby, fbz) &
          bind(C, name="spark_set_flux_corrections")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
       type(c_ptr), intent(in) :: fbx(*), fby(*), fbz(*)
     end function spark_set_flux_corrections

     integer(c_int) function spark_flux_correct(batch, dt) bind(C, name="spark_flux_correct")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: batch
       real(c_double), value :: dt
     end function spark_flux_correct

     integer(c_int) function spark_set_option(batch, option, val) bind(C, name="spark_set_option")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
       integer(c_int), value :: option, val
     end function spark_set_option

     integer(c_int) function spark_batched_tiles(batch, rows, max_rows, nrows) bind(C, name="spark_batched_tiles")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
       integer(c_int), intent(out) :: rows(*)
       integer(c_int), value :: max_rows
       integer(c_int), intent(out) :: nrows
     end function spark_batched_tiles

     integer(c_int) function spark_synchronize(batch) bind(C, name="spark_synchronize")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
     end function spark_synchronize

     integer(c_int) function spark_get_counters(batch, counters, n) bind(C, name="spark_get_counters")
       import :: c_int, c_ptr, c_long_long
       type(c_ptr), value :: batch
       integer(c_long_long), intent(out) :: counters(*)
       integer(c_int), value :: n
     end function spark_get_counters

     integer(c_int) function spark_profile(batch, enable) bind(C, name="spark_profile")
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
       integer(c_int), value :: enable
     end function spark_profile

     integer(c_int) function spark_profile_read(batch, stats, max, n) bind(C, name="spark_profile_read")
       import :: c_int, c_ptr, spark_kernel_stat_t
       type(c_ptr), value :: batch
       type(spark_kernel_stat_t), intent(out) :: stats(*)
       integer(c_int), value :: max
       integer(c_int), intent(out) :: n
     end function spark_profile_read

     integer(c_int) function spark_last_error(buf, len) bind(C, name="spark_last_error")
       import :: c_int, c_char
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_int), value :: len
     end function spark_last_error
  end interface

  abstract interface
     subroutine abort_hook_t(msg)
       character(len=*), intent(in) :: msg
     end subroutine abort_hook_t
  end interface

  procedure(abort_hook_t), pointer :: abort_hook => null()

contains

  !! the library's last error message as a Fortran string
  function spark_error_message() result(msg)
    character(len=512) :: msg
    character(kind=c_char) :: buf(512)
    integer :: i, rc
    rc = spark_last_error(buf, 512_c_int)
    msg = ' '
    do i = 1, 512
       if (buf(i) == c_null_char) exit
       msg(i:i) = buf(i)
    end do
  end function spark_error_message

  !! route failures to a caller-supplied abort routine (Driver_abort in Flash-X)
  subroutine spark_set_abort_hook(hook)
    procedure(abort_hook_t) :: hook
    abort_hook => hook
  end subroutine spark_set_abort_hook

  !! abort with the library's message when status /= SPARK_OK
  subroutine spark_check(status, where)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: where
    character(len=600) :: msg
    if (status == SPARK_OK) return
    write (msg, '(A,A,I0,A,A)') trim(where), ': status ', status, ': ', trim(spark_error_message())
    if (associated(abort_hook)) then
       call abort_hook(trim(msg))
    else
       write (*, '(A)') trim(msg)
       error stop 1
    end if
  end subroutine spark_check

end module spark_hip
