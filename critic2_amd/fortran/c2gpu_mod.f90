! iso_c_binding interfaces of libc2gpu (include/c2gpu.h) for critic2's
! Fortran host.  Conventions follow critic2: grids are f(n1,n2,n3) with x
! fastest, 3x3 matrices are Fortran column-major, lat_i_dist is
! lat_i_dist(-1:1,-1:1,-1:1).  Every function returns 0 on success; the
! caller turns a non-zero status into ferror(...,faterr).
module c2gpu
  use iso_c_binding, only: c_int, c_int32_t, c_int64_t, c_double, c_ptr, c_char, c_funptr, c_null_funptr, &
     c_null_ptr, c_null_char, c_associated, c_f_pointer
  implicit none
  private

  integer(c_int), parameter, public :: C2GPU_OK = 0
  integer(c_int), parameter, public :: C2GPU_PTR_DEVICE = 1
  integer(c_int), parameter, public :: C2GPU_ATTR_NEAREST_ATOM = 0
  integer(c_int), parameter, public :: C2GPU_ATTR_RATOM = 1
  integer(c_int), parameter, public :: C2GPU_ATTR_NOATOMS = 2
  integer(c_int), parameter, public :: C2GPU_FT_GRAD = 0, C2GPU_FT_XX = 1, C2GPU_FT_YY = 2, &
     C2GPU_FT_ZZ = 3, C2GPU_FT_LAP = 4

  type, bind(c), public :: c2gpu_attractor_opts
     integer(c_int) :: mode
     real(c_double) :: ratom
     real(c_double) :: x2c(9)
     integer(c_int) :: nat
     type(c_ptr) :: xat
     ! DISCARD: a bind(c) function  integer(c_int) f(x, user)  with
     ! real(c_double) :: x(3) (Cartesian, bohr) and type(c_ptr), value :: user,
     ! nonzero to discard the new maximum; c_null_funptr for none
     type(c_funptr) :: discard = c_null_funptr
     type(c_ptr) :: discard_user = c_null_ptr
     ! leading entries of xat that are atoms (s%c%ncel); the rest are the
     ! field's other critical points (cpcel order); 0: all are atoms
     integer(c_int) :: nat_atoms = 0
  end type c2gpu_attractor_opts

  type, bind(c), public :: c2gpu_bader_stats
     integer(c_int) :: nattr, nmaxima, sweep_iters, refine_iters
     integer(c_int64_t) :: nstarts, traj_points, edge_points, events_total
     real(c_double) :: ms_traj, ms_sweep, ms_refine, ms_total
  end type c2gpu_bader_stats

  integer(c_int), parameter, public :: C2GPU_CUBE_E22_14 = 0, C2GPU_CUBE_1P_E12_5 = 1, &
     C2GPU_CUBE_1P_E13_5 = 2

  ! header of a cube file (grid3mod@proc.f90:533-551)
  type, bind(c), public :: c2gpu_cube_header
     integer(c_int) :: nat, ismo
     integer(c_int) :: n(3)
     real(c_double) :: x0(3)
     real(c_double) :: step(9)
     integer(c_int64_t) :: data_offset
  end type c2gpu_cube_header

  public :: c2gpu_ctx_create, c2gpu_ctx_destroy, c2gpu_bader_metrics, c2gpu_bader
  public :: c2gpu_basin_sums, c2gpu_fft_deriv, c2gpu_relabel
  public :: c2gpu_cube_read_header, c2gpu_cube_read, c2gpu_bincube_read, c2gpu_cube_write
  public :: c2gpu_nci_cutoff, c2gpu_multipoles, c2gpu_yt_multipoles, c2gpu_yt_remap
  public :: c2gpu_nci, c2gpu_nci_promol_nodes, c2gpu_nci_fragments, c2gpu_nci_cutoff_frag

  integer(c_int), parameter, public :: C2GPU_FP64 = 0, C2GPU_FP32 = 1

  ! nciplot cutoff options (nci@proc.f90:612-643)
  type, bind(c), public :: c2gpu_nci_cut
     real(c_double) :: rhocut, dimcut, rhoplot
     real(c_double) :: srhorange(2)
     integer(c_int) :: onlyneg
     real(c_double) :: rho_void
  end type c2gpu_nci_cut

  integer(c_int), parameter, public :: C2GPU_NCI_FFT = 0, C2GPU_NCI_GRID = 1, C2GPU_NCI_PROMOL = 2

  ! radial density table of a species (grid1, grid1mod.f90:26-48); the
  ! arrays r, f, fp, fpp (ngrid each) by address
  type, bind(c), public :: c2gpu_radial_table
     integer(c_int) :: ngrid
     real(c_double) :: a, b, rmax, cutrad
     type(c_ptr) :: r, f, fp, fpp
  end type c2gpu_radial_table

  ! promolecular density: nspecies tables, nat atoms (fractional xat(3,nat),
  ! species(nat) table indices from 0)
  type, bind(c), public :: c2gpu_promol
     integer(c_int) :: nspecies
     type(c_ptr) :: tables
     integer(c_int) :: nat
     type(c_ptr) :: xat, species
  end type c2gpu_promol

  ! nciplot's nodes x0 + i*xmat(:,1) + j*xmat(:,2) + k*xmat(:,3)
  ! (nci@proc.f90:412-446, :548)
  type, bind(c), public :: c2gpu_nci_nodes
     integer(c_int) :: nstep(3)
     real(c_double) :: x0(3)
     real(c_double) :: xmat(9)
  end type c2gpu_nci_nodes

  interface
     function c2gpu_ctx_create(ctx, device) bind(c, name="c2gpu_ctx_create")
       import :: c_ptr, c_int
       type(c_ptr), intent(out) :: ctx
       integer(c_int), value :: device
       integer(c_int) :: c2gpu_ctx_create
     end function c2gpu_ctx_create

     function c2gpu_ctx_destroy(ctx) bind(c, name="c2gpu_ctx_destroy")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int) :: c2gpu_ctx_destroy
     end function c2gpu_ctx_destroy

     function c2gpu_bader_metrics(x2c, n, lat2car, car2lat, lat_i_dist) &
        bind(c, name="c2gpu_bader_metrics")
       import :: c_int, c_double
       real(c_double), intent(in) :: x2c(9)
       integer(c_int), intent(in) :: n(3)
       real(c_double), intent(out) :: lat2car(9), car2lat(9), lat_i_dist(27)
       integer(c_int) :: c2gpu_bader_metrics
     end function c2gpu_bader_metrics

     function c2gpu_bader(ctx, rho, n, car2lat, lat_i_dist, opts, flags, label_out, nattr, &
        xattr_out, xattr_cap, stats) bind(c, name="c2gpu_bader")
       import :: c_ptr, c_int, c_int32_t, c_double, c2gpu_attractor_opts, c2gpu_bader_stats
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: rho(*)
       integer(c_int), intent(in) :: n(3)
       real(c_double), intent(in) :: car2lat(9), lat_i_dist(27)
       type(c2gpu_attractor_opts), intent(in) :: opts
       integer(c_int), value :: flags
       integer(c_int32_t), intent(out) :: label_out(*)
       integer(c_int), intent(out) :: nattr
       real(c_double), intent(out) :: xattr_out(*)
       integer(c_int), value :: xattr_cap
       type(c2gpu_bader_stats), intent(out) :: stats
       integer(c_int) :: c2gpu_bader
     end function c2gpu_bader

     function c2gpu_relabel(ctx, npts, map, nmap, flags, label) bind(c, name="c2gpu_relabel")
       import :: c_ptr, c_int, c_int32_t, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: npts
       integer(c_int32_t), intent(in) :: map(*)
       integer(c_int), value :: nmap, flags
       integer(c_int32_t), intent(inout) :: label(*)
       integer(c_int) :: c2gpu_relabel
     end function c2gpu_relabel

     function c2gpu_basin_sums(ctx, label, npts, nattr, nfields, fields, flags, sums_out, &
        counts_out) bind(c, name="c2gpu_basin_sums")
       import :: c_ptr, c_int, c_int32_t, c_int64_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: label(*)
       integer(c_int64_t), value :: npts
       integer(c_int), value :: nattr, nfields
       type(c_ptr), intent(in) :: fields(*)
       integer(c_int), value :: flags
       real(c_double), intent(out) :: sums_out(*)
       integer(c_int64_t), intent(out) :: counts_out(*)
       integer(c_int) :: c2gpu_basin_sums
     end function c2gpu_basin_sums

     function c2gpu_fft_deriv(ctx, rho, n, x2c, kind, flags, out) bind(c, name="c2gpu_fft_deriv")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: rho(*)
       integer(c_int), intent(in) :: n(3)
       real(c_double), intent(in) :: x2c(9)
       integer(c_int), value :: kind, flags
       real(c_double), intent(out) :: out(*)
       integer(c_int) :: c2gpu_fft_deriv
     end function c2gpu_fft_deriv

     function c2gpu_cube_read_header(ctx, path, binary, hdr) bind(c, name="c2gpu_cube_read_header")
       import :: c_ptr, c_int, c_char, c2gpu_cube_header
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int), value :: binary
       type(c2gpu_cube_header), intent(out) :: hdr
       integer(c_int) :: c2gpu_cube_read_header
     end function c2gpu_cube_read_header

     function c2gpu_cube_read(ctx, path, flags, hdr, f) bind(c, name="c2gpu_cube_read")
       import :: c_ptr, c_int, c_char, c_double, c2gpu_cube_header
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int), value :: flags
       type(c2gpu_cube_header), intent(out) :: hdr
       real(c_double), intent(out) :: f(*)
       integer(c_int) :: c2gpu_cube_read
     end function c2gpu_cube_read

     function c2gpu_bincube_read(ctx, path, flags, hdr, f) bind(c, name="c2gpu_bincube_read")
       import :: c_ptr, c_int, c_char, c_double, c2gpu_cube_header
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int), value :: flags
       type(c2gpu_cube_header), intent(out) :: hdr
       real(c_double), intent(out) :: f(*)
       integer(c_int) :: c2gpu_bincube_read
     end function c2gpu_bincube_read

     function c2gpu_cube_write(ctx, path, hdr, zat, xat, f, style, flags) bind(c, name="c2gpu_cube_write")
       import :: c_ptr, c_int, c_char, c_double, c2gpu_cube_header
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
       type(c2gpu_cube_header), intent(in) :: hdr
       integer(c_int), intent(in) :: zat(*)
       real(c_double), intent(in) :: xat(*)
       real(c_double), intent(in) :: f(*)
       integer(c_int), value :: style, flags
       integer(c_int) :: c2gpu_cube_write
     end function c2gpu_cube_write

     function c2gpu_nci_cutoff(ctx, crho, cgrad, rhoat, n, precision, opts, flags, dat, cap, len, sums) &
        bind(c, name="c2gpu_nci_cutoff")
       import :: c_ptr, c_int, c_int64_t, c_double, c_char, c2gpu_nci_cut
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: crho(*)
       real(c_double), intent(inout) :: cgrad(*)
       real(c_double), intent(in) :: rhoat(*)
       integer(c_int), intent(in) :: n(3)
       integer(c_int), value :: precision
       type(c2gpu_nci_cut), intent(in) :: opts
       integer(c_int), value :: flags
       character(kind=c_char), intent(out) :: dat(*)
       integer(c_int64_t), value :: cap
       integer(c_int64_t), intent(out) :: len
       real(c_double), intent(out) :: sums(3)
       integer(c_int) :: c2gpu_nci_cutoff
     end function c2gpu_nci_cutoff

     function c2gpu_nci(ctx, rho, n, x2c, scheme, precision, promol, zfast, flags, crho_out, cgrad_out) &
        bind(c, name="c2gpu_nci")
       import :: c_ptr, c_int, c_double, c2gpu_promol
       type(c_ptr), value :: ctx
       type(c_ptr), value :: rho
       integer(c_int), intent(in) :: n(3)
       real(c_double), intent(in) :: x2c(9)
       integer(c_int), value :: scheme, precision
       type(c2gpu_promol), intent(in) :: promol
       integer(c_int), value :: zfast, flags
       real(c_double), intent(out) :: crho_out(*), cgrad_out(*)
       integer(c_int) :: c2gpu_nci
     end function c2gpu_nci

     function c2gpu_nci_promol_nodes(ctx, nodes, x2c, promol, precision, flags, crho_out, cgrad_out, &
        rhoat_out) bind(c, name="c2gpu_nci_promol_nodes")
       import :: c_ptr, c_int, c_double, c2gpu_promol, c2gpu_nci_nodes
       type(c_ptr), value :: ctx
       type(c2gpu_nci_nodes), intent(in) :: nodes
       real(c_double), intent(in) :: x2c(9)
       type(c2gpu_promol), intent(in) :: promol
       integer(c_int), value :: precision, flags
       type(c_ptr), value :: crho_out, cgrad_out, rhoat_out
       integer(c_int) :: c2gpu_nci_promol_nodes
     end function c2gpu_nci_promol_nodes

     function c2gpu_nci_fragments(ctx, nodes, nspecies, tables, nfrag, frag_nat, xyz, species, precision, &
        flags, rhofrag_out) bind(c, name="c2gpu_nci_fragments")
       import :: c_ptr, c_int, c_double, c2gpu_nci_nodes, c2gpu_radial_table
       type(c_ptr), value :: ctx
       type(c2gpu_nci_nodes), intent(in) :: nodes
       integer(c_int), value :: nspecies
       type(c2gpu_radial_table), intent(in) :: tables(*)
       integer(c_int), value :: nfrag
       integer(c_int), intent(in) :: frag_nat(*)
       real(c_double), intent(in) :: xyz(*)
       integer(c_int), intent(in) :: species(*)
       integer(c_int), value :: precision, flags
       type(c_ptr), value :: rhofrag_out
       integer(c_int) :: c2gpu_nci_fragments
     end function c2gpu_nci_fragments

     function c2gpu_nci_cutoff_frag(ctx, crho, cgrad, rhoat, rhofrag, nfrag, rhoparam, rhoparam2, n, &
        precision, opts, flags, dat, cap, len, sums) bind(c, name="c2gpu_nci_cutoff_frag")
       import :: c_ptr, c_int, c_int64_t, c_double, c_char, c2gpu_nci_cut
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: crho(*)
       real(c_double), intent(inout) :: cgrad(*)
       real(c_double), intent(in) :: rhoat(*)
       real(c_double), intent(in) :: rhofrag(*)
       integer(c_int), value :: nfrag
       real(c_double), value :: rhoparam, rhoparam2
       integer(c_int), intent(in) :: n(3)
       integer(c_int), value :: precision
       type(c2gpu_nci_cut), intent(in) :: opts
       integer(c_int), value :: flags
       character(kind=c_char), intent(out) :: dat(*)
       integer(c_int64_t), value :: cap
       integer(c_int64_t), intent(out) :: len
       real(c_double), intent(out) :: sums(3)
       integer(c_int) :: c2gpu_nci_cutoff_frag
     end function c2gpu_nci_cutoff_frag

     function c2gpu_multipoles(ctx, label, f, n, x2c, nattr, xattr, lmax, flags, mpole_out) &
        bind(c, name="c2gpu_multipoles")
       import :: c_ptr, c_int, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: label(*)
       real(c_double), intent(in) :: f(*)
       integer(c_int), intent(in) :: n(3)
       real(c_double), intent(in) :: x2c(9)
       integer(c_int), value :: nattr
       real(c_double), intent(in) :: xattr(*)
       integer(c_int), value :: lmax, flags
       real(c_double), intent(out) :: mpole_out(*)
       integer(c_int) :: c2gpu_multipoles
     end function c2gpu_multipoles

     function c2gpu_yt_multipoles(ctx, label, f, n, x2c, nattr, xattr, lmax, flags, mpole_out) &
        bind(c, name="c2gpu_yt_multipoles")
       import :: c_ptr, c_int, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: label(*)
       real(c_double), intent(in) :: f(*)
       integer(c_int), intent(in) :: n(3)
       real(c_double), intent(in) :: x2c(9)
       integer(c_int), value :: nattr
       real(c_double), intent(in) :: xattr(*)
       integer(c_int), value :: lmax, flags
       real(c_double), intent(out) :: mpole_out(*)
       integer(c_int) :: c2gpu_yt_multipoles
     end function c2gpu_yt_multipoles

     function c2gpu_yt_remap(ctx, map, nmap, nattr_new) bind(c, name="c2gpu_yt_remap")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: map(*)
       integer(c_int), value :: nmap, nattr_new
       integer(c_int) :: c2gpu_yt_remap
     end function c2gpu_yt_remap

     function c2gpu_ctx_error(ctx) bind(c, name="c2gpu_ctx_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
       type(c_ptr) :: c2gpu_ctx_error
     end function c2gpu_ctx_error
  end interface

  public :: c2gpu_error_message

contains

  ! The context's last error message (c2gpu_ctx_error) as a Fortran string,
  ! for ferror.
  subroutine c2gpu_error_message(ctx, msg)
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(out) :: msg
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: i
    msg = ""
    p = c2gpu_ctx_error(ctx)
    if (.not.c_associated(p)) return
    call c_f_pointer(p, cs, (/len(msg)/))
    do i = 1, len(msg)
       if (cs(i) == c_null_char) exit
       msg(i:i) = cs(i)
    end do
  end subroutine c2gpu_error_message

end module c2gpu
