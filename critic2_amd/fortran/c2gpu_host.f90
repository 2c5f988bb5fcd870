! Stand-in for critic2's Fortran host on this boundary: reads a grid and a
! cell in critic2's layout, calls the drop-in through the c2gpu module
! exactly as a bader@c2gpu submodule would (metrics, labelling, per-basin
! Volume and Pop with the Omega/N factor applied on the Fortran side,
! integration@proc.f90:1215 and :1296), and writes the labels and the
! integrated properties.
!
! Input file (text): n1 n2 n3 / x2c(3,3) column by column / nat / xat(3,i)
! per atom, then the grid as an unformatted stream file next to it, or as a
! Gaussian cube (a name ending in .cube) read through c2gpu_cube_read, the
! drop-in for grid3%read_cube.
! Usage: c2gpu_host <case.txt> <grid.bin|grid.cube> <labels.bin> <props.txt>
!                   [mode ratom [discard_below_x]]
! mode: 0 nearest atom (default), 1 NNM/RATOM, 2 NOATOMS; with a third value
! the DISCARD callback rejects every new maximum whose Cartesian x is below
! it.  The attractor array starts at nat entries and is grown from the
! count the library reports, as the bader@c2gpu submodule does.
module host_discard
  use iso_c_binding, only: c_int, c_double, c_ptr
  implicit none
  real(c_double) :: xlimit = 0d0
contains
  ! DISCARD as a bind(c) callback: nonzero discards the maximum at x (bohr)
  function discard_below(x, user) result(res) bind(c)
    real(c_double), intent(in) :: x(3)
    type(c_ptr), value :: user
    integer(c_int) :: res
    res = 0
    if (x(1) < xlimit) res = 1
  end function discard_below
end module host_discard

program c2gpu_host
  use iso_c_binding, only: c_ptr, c_int, c_int32_t, c_int64_t, c_double, c_loc, c_null_ptr, c_null_char, &
     c_funloc
  use c2gpu
  use host_discard
  implicit none
  character(len=512) :: fcase, fgrid, flab, fprop, arg
  integer :: mode, cap
  real(c_double) :: ratom
  integer(c_int) :: n(3), nat, ier, nattr
  real(c_double) :: x2c(9), lat2car(9), car2lat(9), lat_i_dist(27), omega
  real(c_double), allocatable, target :: f(:,:,:), xat(:,:), xattr(:,:), sums(:)
  integer(c_int32_t), allocatable :: idg(:,:,:)
  integer(c_int64_t), allocatable :: counts(:)
  type(c_ptr) :: ctx
  type(c_ptr) :: fields(1)
  type(c2gpu_attractor_opts) :: opts
  type(c2gpu_bader_stats) :: stats
  type(c2gpu_cube_header) :: hdr
  integer :: lu, i, lg
  integer(c_int64_t) :: npts

  call get_command_argument(1, fcase)
  call get_command_argument(2, fgrid)
  call get_command_argument(3, flab)
  call get_command_argument(4, fprop)
  mode = C2GPU_ATTR_NEAREST_ATOM
  ratom = 1d40
  if (command_argument_count() >= 6) then
     call get_command_argument(5, arg)
     read(arg,*) mode
     call get_command_argument(6, arg)
     read(arg,*) ratom
  end if

  open(newunit=lu, file=trim(fcase), status='old')
  read(lu,*) n
  read(lu,*) x2c
  read(lu,*) nat
  allocate(xat(3,nat))
  read(lu,*) xat
  close(lu)
  allocate(f(n(1),n(2),n(3)), idg(n(1),n(2),n(3)))
  ier = c2gpu_ctx_create(ctx, 0_c_int)
  if (ier /= C2GPU_OK) call die("c2gpu_ctx_create", ier)
  lg = len_trim(fgrid)
  if (lg > 5 .and. fgrid(max(lg-4,1):lg) == '.cube') then
     ier = c2gpu_cube_read(ctx, trim(fgrid) // c_null_char, 0_c_int, hdr, f)
     if (ier /= C2GPU_OK) call die("c2gpu_cube_read", ier)
     if (any(hdr%n /= n)) call die("cube grid size", 1_c_int)
  else
     open(newunit=lu, file=trim(fgrid), access='stream', form='unformatted', status='old')
     read(lu) f
     close(lu)
  end if
  ier = c2gpu_bader_metrics(x2c, n, lat2car, car2lat, lat_i_dist)
  if (ier /= C2GPU_OK) call die("c2gpu_bader_metrics", ier)

  opts%mode = mode
  opts%ratom = ratom
  opts%x2c = x2c
  opts%nat = nat
  opts%xat = c_loc(xat)
  if (command_argument_count() >= 7) then
     call get_command_argument(7, arg)
     read(arg,*) xlimit
     opts%discard = c_funloc(discard_below)
  end if
  ! the attractor array starts at nat entries; a larger count is read back
  ! and the call repeated with room for every attractor
  cap = max(nat,1)
  do
     if (allocated(xattr)) deallocate(xattr)
     allocate(xattr(3,cap))
     ier = c2gpu_bader(ctx, f, n, car2lat, lat_i_dist, opts, 0_c_int, idg, nattr, xattr, cap, stats)
     if (ier /= C2GPU_OK) call die("c2gpu_bader", ier)
     if (nattr <= cap) exit
     cap = nattr
  end do

  npts = int(n(1),c_int64_t) * n(2) * n(3)
  allocate(sums(max(nattr,1)), counts(max(nattr,1)))
  fields(1) = c_loc(f)
  ier = c2gpu_basin_sums(ctx, idg, npts, nattr, 1_c_int, fields, 0_c_int, sums, counts)
  if (ier /= C2GPU_OK) call die("c2gpu_basin_sums", ier)
  ier = c2gpu_ctx_destroy(ctx)

  omega = abs(x2c(1)*(x2c(5)*x2c(9)-x2c(6)*x2c(8)) - x2c(4)*(x2c(2)*x2c(9)-x2c(3)*x2c(8)) &
     + x2c(7)*(x2c(2)*x2c(6)-x2c(3)*x2c(5)))
  open(newunit=lu, file=trim(flab), access='stream', form='unformatted', status='replace')
  write(lu) idg
  close(lu)
  open(newunit=lu, file=trim(fprop), status='replace')
  write(lu,'("nattr ",I0," sweep_iters ",I0," refine_iters ",I0," calls_cap ",I0)') nattr, stats%sweep_iters, &
     stats%refine_iters, cap
  do i = 1, nattr
     write(lu,'(I6,2(1X,ES24.16),1X,I12,3(1X,ES24.16))') i, real(counts(i),8) * omega / real(npts,8), &
        sums(i) * omega / real(npts,8), counts(i), xattr(:,i)
  end do
  close(lu)

contains
  subroutine die(what, code)
    character(*), intent(in) :: what
    integer(c_int), intent(in) :: code
    write(*,'("ERROR ",A," status ",I0)') what, code
    stop 1
  end subroutine die
end program c2gpu_host
