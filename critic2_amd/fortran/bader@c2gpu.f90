! The drop-in for critic2's Bader implementation: a submodule of critic2's
! `bader` module (interface bader.f90:47-61) that a critic2 build compiles
! *instead of* bader@proc.f90.  integration@proc.f90, bader.f90 and the
! keyword surface stay unchanged.  It needs critic2's own modules (systemmod,
! types, tools_io, tools_math, param) and the c2gpu binding module
! (c2gpu_mod.f90), so it is compiled inside a critic2 build, not by this
! package's build(); the stand-in host c2gpu_host.f90 drives the same calls.
!
!  bader_integrate  bader@proc.f90:80-234: metrics with critic2's own matinv,
!                   labels and attractors from c2gpu_bader (all attractor
!                   rules, DISCARD through a callback that evaluates the
!                   expression with critic2's s%eval), bas%xattr sized to
!                   what the library reports.
!  bader_remap      bader@proc.f90:237-296: the attractor + lattice-vector
!                   split of each basin, on the returned labels, with
!                   critic2's own crystal%shortest.
submodule (bader) c2gpu_impl
  use iso_c_binding
  use c2gpu
  use systemmod, only: system
  implicit none

  ! the system and expression the DISCARD callback evaluates (set for the
  ! duration of one bader_integrate call)
  type(system), pointer :: disc_sys => null()
  character(len=:), allocatable :: disc_expr

contains

  ! DISCARD (bader@proc.f90:185-191): nonzero discards the maximum at the
  ! Cartesian position x (bohr).
  function discard_cb(x, user) result(res) bind(c)
    use tools_io, only: ferror, faterr
    real(c_double), intent(in) :: x(3)
    type(c_ptr), value :: user
    integer(c_int) :: res
    real*8 :: fval
    character(len=:), allocatable :: errmsg

    fval = disc_sys%eval(disc_expr, errmsg, x)
    if (len_trim(errmsg) > 0) &
       call ferror("bader_integrate","invalid DISCARD expression: " // trim(errmsg),faterr)
    res = 0
    if (.not.(abs(fval) < 1d-30)) res = 1
  end function discard_cb

  module subroutine bader_integrate(s,bas,iref)
    use systemmod, only: system
    use tools_io, only: ferror, faterr, string
    use tools_math, only: matinv
    use types, only: basindat, realloc
    type(system), intent(inout), target :: s
    type(basindat), intent(inout) :: bas
    integer, intent(in) :: iref

    type(c_ptr) :: ctx
    type(c2gpu_attractor_opts) :: opts
    type(c2gpu_bader_stats) :: stats
    real(c_double) :: lat2car(3,3), car2lat(3,3), lid(-1:1,-1:1,-1:1), dcar(3)
    real(c_double), allocatable, target :: xat(:,:)
    integer(c_int) :: n(3), ier, nattr, cap
    integer :: i, j, k
    character(len=1024) :: msg

    n = shape(bas%f)
    ! grid metrics with critic2's own matinv (LAPACK), bader@proc.f90:121-145
    do i = 1, 3
       lat2car(:,i) = s%c%m_x2c(:,i) / n(i)
    end do
    car2lat = lat2car
    call matinv(car2lat,3)
    lid = 0d0
    do i = -1, 1
       do j = -1, 1
          do k = -1, 1
             if (i == 0 .and. j == 0 .and. k == 0) cycle
             dcar = matmul(lat2car,real((/i,j,k/),8))
             lid(i,j,k) = 1d0 / sqrt(sum(dcar*dcar))
          end do
       end do
    end do

    ! attractor options as intgrid_driver sets them (integration@proc.f90:236-248);
    ! the known attractors are the field's critical points, atoms first
    ! (bader@proc.f90:109-117), and only the atoms are identify_atom's
    allocate(xat(3,max(s%f(iref)%ncpcel,1)))
    xat = 0d0
    do i = 1, s%f(iref)%ncpcel
       xat(:,i) = s%f(iref)%cpcel(i)%x
    end do
    opts%mode = C2GPU_ATTR_NEAREST_ATOM
    if (.not.bas%atexist) then
       opts%mode = C2GPU_ATTR_NOATOMS
    else if (bas%ratom < 1d39) then
       opts%mode = C2GPU_ATTR_RATOM
    end if
    opts%ratom = bas%ratom
    opts%x2c = reshape(s%c%m_x2c,(/9/))
    opts%nat = s%f(iref)%ncpcel
    opts%nat_atoms = s%c%ncel
    opts%xat = c_loc(xat)
    if (len_trim(bas%expr) > 0) then
       disc_sys => s
       disc_expr = trim(bas%expr)
       opts%discard = c_funloc(discard_cb)
    end if

    ier = c2gpu_ctx_create(ctx, 0_c_int)
    if (ier /= C2GPU_OK) call ferror("bader_integrate","c2gpu_ctx_create failed",faterr)
    if (allocated(bas%idg)) deallocate(bas%idg)
    allocate(bas%idg(n(1),n(2),n(3)))
    ! room for the atoms and the non-nuclear maxima of NNM/RATOM/NOATOMS; the
    ! library reports the full count, so a short array is grown and the call
    ! repeated
    cap = 4 * max(opts%nat,1) + 4096
    do
       if (allocated(bas%xattr)) deallocate(bas%xattr)
       allocate(bas%xattr(3,cap))
       ier = c2gpu_bader(ctx, bas%f, n, reshape(car2lat,(/9/)), reshape(lid,(/27/)), opts, &
          0_c_int, bas%idg, nattr, bas%xattr, cap, stats)
       if (ier /= C2GPU_OK) exit
       if (nattr <= cap) exit
       cap = nattr
    end do
    if (ier /= C2GPU_OK) then
       call c2gpu_error_message(ctx, msg)
       call ferror("bader_integrate","GPU Bader failed (status " // string(ier) // "): " // trim(msg),faterr)
    end if
    call realloc(bas%xattr,3,max(nattr,1))
    bas%nattr = nattr
    ier = c2gpu_ctx_destroy(ctx)
    disc_sys => null()
  end subroutine bader_integrate

  ! bader_remap (bader@proc.f90:237-296): every grid point p of basin
  ! a = idg(p) lies at x = (p-1)/n - xattr(:,a) from its attractor; the
  ! lattice vector l that crystal%shortest removes from x names the image of
  ! the attractor the point belongs to.  Points with l = 0 keep a; each new
  ! pair (a, l) becomes attractor nattr+1, nattr+2, ... in the order the
  ! grid (x fastest) first meets it; idg1 holds the new ids, ilvec(:,m) the
  ! lattice vector and iatt(m) the original attractor of id m.
  module subroutine bader_remap(s,bas,nattn,idg1,ilvec,iatt)
    use types, only: realloc, basindat
    type(system), intent(in) :: s
    type(basindat), intent(in) :: bas
    integer, intent(out) :: nattn
    integer, allocatable, intent(inout) :: iatt(:), ilvec(:,:), idg1(:,:,:)

    integer :: m1, m2, m3, a, id, l(3), q
    real*8 :: x(3), xs(3), dist

    nattn = bas%nattr
    if (allocated(iatt)) deallocate(iatt)
    if (allocated(ilvec)) deallocate(ilvec)
    if (allocated(idg1)) deallocate(idg1)
    allocate(iatt(max(2*bas%nattr,16)), ilvec(3,max(2*bas%nattr,16)))
    allocate(idg1(bas%n(1),bas%n(2),bas%n(3)))
    ilvec = 0
    do q = 1, bas%nattr
       iatt(q) = q
    end do
    do m3 = 1, bas%n(3)
       do m2 = 1, bas%n(2)
          do m1 = 1, bas%n(1)
             a = bas%idg(m1,m2,m3)
             idg1(m1,m2,m3) = a
             if (a < 1) cycle  ! a basin DISCARD left unlabelled stays 0
             x = real((/m1,m2,m3/) - 1,8) / bas%n - bas%xattr(:,a)
             xs = x
             call s%c%shortest(xs,dist)
             l = nint(x - s%c%c2x(xs))
             if (all(l == 0)) cycle
             ! an image met before keeps its id
             id = 0
             do q = bas%nattr+1, nattn
                if (iatt(q) == a .and. all(ilvec(:,q) == l)) then
                   id = q
                   exit
                end if
             end do
             if (id == 0) then
                nattn = nattn + 1
                if (nattn > size(iatt)) then
                   call realloc(iatt,2*nattn)
                   call realloc(ilvec,3,2*nattn)
                end if
                iatt(nattn) = a
                ilvec(:,nattn) = l
                id = nattn
             end if
             idg1(m1,m2,m3) = id
          end do
       end do
    end do
    call realloc(iatt,nattn)
    call realloc(ilvec,3,nattn)
  end subroutine bader_remap

end submodule c2gpu_impl
