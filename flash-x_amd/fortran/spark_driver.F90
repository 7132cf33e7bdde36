!! spark_driver.F90 -- a Fortran driver of libspark_hip.so shaped like the
!! Flash-X unsplit driver loop (Flash-X/source/Driver/DriverMain/Unsplit/
!! Driver_evolveAll.F90: Driver_computeDt -> Hydro per step).
!!
!! The block data live in a Flash-X style array
!!     unk(NUNK_VARS, ILO_GC:IHI_GC, JLO_GC:JHI_GC, KLO_GC:KHI_GC, MAXBLOCKS)
!! (Paramesh4/PM4_package/headers/physicaldata.F90:80), variable index
!! fastest, and are handed to the library as one base address per block,
!! c_loc(unk(1, ILO_GC, JLO_GC, KLO_GC, blockID)), exactly what
!! tileDesc%getDataPtr exposes to Spark (Grid/GridMain/AMR/Paramesh4/
!! Grid_tile.F90:230).
!!
!! Usage:  spark_driver <problem file> <result file>
!!
!! Problem file (native-endian stream):
!!   int32  ndim, nx, ny, nz, nbx, nby, nbz, nguard, nunk, nsteps,
!!          guard_mode, nscalars, recon, riemann, rk_order, nstencil,
!!          use_flattening, form, roundtrip
!!   real64 dx, dy, dz, gamma, cfl
!!   real64 unk(nunk, NXg, NYg, NZg, nbx*nby*nbz)   (block b = bi + nbx*(bj + nby*bk))
!! Result file:
!!   int32  nsteps, status;  real64 dt(nsteps);  real64 unk(...) after the run
!!
!! roundtrip = 1 packs unk before and unpacks it after every step (the
!! strict drop-in, host unk authoritative between Hydro calls); 0 keeps the
!! state on the device for the whole run.
module hydro_hip
  use iso_c_binding
  use spark_hip
  implicit none

  type(c_ptr), save :: batch = c_null_ptr
  type(c_ptr), allocatable, save :: blockPtrs(:)

contains

  !! Hydro_init (Spark/Hydro_init.F90:24-179): runtime parameters -> spark_params,
  !! device buffers for every block of the run.
  subroutine Hydro_init(p, blocks, nbx, nby, nbz)
    type(spark_params_t), intent(in) :: p
    type(spark_block_desc_t), intent(in) :: blocks(:)
    integer, intent(in) :: nbx, nby, nbz
    call spark_check(spark_init(0_c_int), 'spark_init')
    call spark_check(spark_batch_create(p, int(size(blocks), c_int), blocks, batch), 'spark_batch_create')
    call spark_check(spark_set_uniform_grid(batch, int(nbx, c_int), int(nby, c_int), int(nbz, c_int)), &
                     'spark_set_uniform_grid')
  end subroutine Hydro_init

  !! Hydro_computeDt (Hydro_interface.F90:24-42), reduced over all blocks of
  !! the batch in one call; dt_minloc = (i, j, k, level, block): i, j, k
  !! 1-based in the guard-celled box, block the 0-based batch index.
  subroutine Hydro_computeDt(dt_check, dt_minloc)
    real(c_double), intent(out) :: dt_check
    integer(c_int), intent(out) :: dt_minloc(5)
    call spark_check(spark_compute_dt(batch, dt_check, dt_minloc), 'spark_compute_dt')
  end subroutine Hydro_computeDt

  !! Hydro (Hydro_interface.F90:61-66): one RK step of every block.
  subroutine Hydro(timeEndAdv, dt, dtOld, sweepOrder)
    real, intent(in) :: timeEndAdv, dt, dtOld
    integer, intent(in) :: sweepOrder
    call spark_check(spark_hydro_step(batch, real(dt, c_double), real(dtOld, c_double)), 'spark_hydro_step')
  end subroutine Hydro

  subroutine Hydro_finalize()
    call spark_check(spark_batch_destroy(batch), 'spark_batch_destroy')
    batch = c_null_ptr
    call spark_check(spark_finalize(), 'spark_finalize')
  end subroutine Hydro_finalize

end module hydro_hip

program spark_driver
  use iso_c_binding
  use spark_hip
  use hydro_hip
  implicit none

  integer(c_int) :: hdr(19), dt_minloc(5), status
  real(c_double) :: rhdr(5)
  integer :: ndim, nx, ny, nz, nbx, nby, nbz, ng, nunk, nsteps, nblk, roundtrip
  integer :: ilo, ihi, jlo, jhi, klo, khi, b, step, u
  real(c_double), allocatable, target :: unk(:, :, :, :, :)
  real(c_double), allocatable :: dts(:)
  type(spark_params_t) :: p
  type(spark_block_desc_t), allocatable :: blocks(:)
  real(c_double) :: dt, dtOld
  character(len=1024) :: inFile, outFile

  call get_command_argument(1, inFile)
  call get_command_argument(2, outFile)
  open (newunit=u, file=trim(inFile), access='stream', form='unformatted', status='old', action='read')
  read (u) hdr, rhdr
  ndim = hdr(1); nx = hdr(2); ny = hdr(3); nz = hdr(4)
  nbx = hdr(5); nby = hdr(6); nbz = hdr(7); ng = hdr(8); nunk = hdr(9); nsteps = hdr(10)
  roundtrip = hdr(19)
  nblk = nbx * nby * nbz

  ! GC box bounds as Flash-X declares them (interior starts at NGUARD+1)
  ilo = 1; ihi = nx + 2 * ng
  jlo = 1; jhi = 1; klo = 1; khi = 1
  if (ndim > 1) jhi = ny + 2 * ng
  if (ndim > 2) khi = nz + 2 * ng
  allocate (unk(nunk, ilo:ihi, jlo:jhi, klo:khi, nblk))
  read (u) unk
  close (u)

  call spark_check(spark_params_default(p), 'spark_params_default')
  p%ndim = ndim; p%nguard = ng; p%nunk_vars = nunk
  p%guard_mode = hdr(11); p%nscalars = hdr(12); p%recon = hdr(13); p%riemann = hdr(14)
  p%rk_order = hdr(15); p%nstencil = hdr(16); p%use_flattening = hdr(17); p%form = hdr(18)
  p%gamma = rhdr(4); p%cfl = rhdr(5)

  allocate (blocks(nblk), blockPtrs(nblk), dts(max(1, nsteps)))
  do b = 1, nblk
     blocks(b) = spark_block_desc_t(nx, ny, nz, 1, rhdr(1), rhdr(2), rhdr(3))
     blockPtrs(b) = c_loc(unk(1, ilo, jlo, klo, b))
  end do

  call Hydro_init(p, blocks, nbx, nby, nbz)
  call spark_check(spark_pack(batch, blockPtrs), 'spark_pack')
  dtOld = 0.0d0
  do step = 1, nsteps
     if (roundtrip == 1 .and. step > 1) call spark_check(spark_pack(batch, blockPtrs), 'spark_pack')
     call Hydro_computeDt(dt, dt_minloc)
     call Hydro(0.0, real(dt), real(dtOld), 0)
     dts(step) = dt
     dtOld = dt
     if (roundtrip == 1) call spark_check(spark_unpack(batch, blockPtrs, 1_c_int, 0_c_int), 'spark_unpack')
  end do
  if (roundtrip /= 1) call spark_check(spark_unpack(batch, blockPtrs, 1_c_int, 0_c_int), 'spark_unpack')
  call Hydro_finalize()

  status = 0
  open (newunit=u, file=trim(outFile), access='stream', form='unformatted', status='replace', action='write')
  write (u) int(nsteps, c_int), status, dts(1:nsteps), unk
  close (u)
  write (*, '(A,I0,A,I0,A,ES12.5)') 'spark_driver: ', nblk, ' blocks, ', nsteps, ' steps, last dt ', dts(max(1, nsteps))
end program spark_driver
