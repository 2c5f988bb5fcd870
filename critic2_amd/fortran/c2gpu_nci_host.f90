! Stand-in for critic2's nciplot on this boundary (nci@proc.f90:525-661 with
! the promolecular field and FRAGMENT input): radial tables, cell, atoms,
! the node set, fragments and cutoff options in, through the c2gpu module
! as an interposed nciplot block would call it: the promolecular block with
! rhoat on the nodes, the fragment densities, and the cutoff pass with the
! fragment test (the .dat rows, the masked RDG, the void sums).
!
! Input (stream file, native byte order):
!   nspecies; per species ngrid, a, b, rmax, cutrad, r(ngrid), f, fp, fpp
!   x2c(9); nat; xat(3,nat); species(nat) (table index from 0)
!   nstep(3); x0(3); xmat(9)
!   nfrag; frag_nat(nfrag); xyz(3,sum(frag_nat)); fspecies(sum(frag_nat))
!   rhocut, dimcut, rhoplot, srhorange(2); onlyneg; rho_void, rhoparam, rhoparam2
! Output (stream file): len (int64), the .dat text (len bytes), crho, cgrad
! (masked), rhoat, rhofrag (z fastest, FP64), sums(3).
! Usage: c2gpu_nci_host <in.bin> <out.bin>
program c2gpu_nci_host
  use iso_c_binding, only: c_ptr, c_int, c_int64_t, c_double, c_char, c_loc
  use c2gpu
  implicit none
  character(len=512) :: fin, fout
  integer(c_int) :: ns, nat, nfrag, ntot, ier, n(3), onlyneg, s, tot
  integer(c_int), allocatable :: ngrid(:), toff(:), frag_nat(:), fspecies(:)
  real(c_double), allocatable :: ta(:), tb(:), trmax(:), tcut(:)
  real(c_double), allocatable, target :: r(:), f(:), fp(:), fpp(:), xat(:), crho(:), cgrad(:), rhoat(:), &
     rhofrag(:)
  integer(c_int), allocatable, target :: spc_t(:)
  real(c_double), allocatable :: xyz(:)
  type(c2gpu_radial_table), allocatable, target :: tables(:)
  type(c2gpu_promol) :: pm
  type(c2gpu_nci_nodes) :: nodes
  type(c2gpu_nci_cut) :: opts
  real(c_double) :: x2c(9), rhoparam, rhoparam2, sums(3)
  character(kind=c_char), allocatable :: dat(:)
  integer(c_int64_t) :: npts, dlen
  type(c_ptr) :: ctx
  integer :: lu

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=lu, file=trim(fin), access='stream', form='unformatted', status='old')
  read(lu) ns
  allocate(ngrid(ns), toff(ns), ta(ns), tb(ns), trmax(ns), tcut(ns), tables(ns))
  ! the tables, one after the other, into concatenated arrays
  allocate(r(0), f(0), fp(0), fpp(0))
  tot = 0
  do s = 1, ns
     read(lu) ngrid(s), ta(s), tb(s), trmax(s), tcut(s)
     toff(s) = tot
     call grow(r, ngrid(s))
     call grow(f, ngrid(s))
     call grow(fp, ngrid(s))
     call grow(fpp, ngrid(s))
     read(lu) r(tot+1:tot+ngrid(s)), f(tot+1:tot+ngrid(s)), fp(tot+1:tot+ngrid(s)), fpp(tot+1:tot+ngrid(s))
     tot = tot + ngrid(s)
  end do
  read(lu) x2c
  read(lu) nat
  allocate(xat(3*nat), spc_t(nat))
  read(lu) xat, spc_t
  read(lu) nodes%nstep, nodes%x0, nodes%xmat
  read(lu) nfrag
  allocate(frag_nat(nfrag))
  read(lu) frag_nat
  ntot = sum(frag_nat)
  allocate(xyz(3*max(ntot,1)), fspecies(max(ntot,1)))
  if (ntot > 0) read(lu) xyz(1:3*ntot), fspecies(1:ntot)
  read(lu) opts%rhocut, opts%dimcut, opts%rhoplot, opts%srhorange, onlyneg, opts%rho_void, rhoparam, rhoparam2
  opts%onlyneg = onlyneg
  close(lu)

  ! tables point into the concatenated arrays (grid1 of each species)
  do s = 1, ns
     tables(s)%ngrid = ngrid(s)
     tables(s)%a = ta(s)
     tables(s)%b = tb(s)
     tables(s)%rmax = trmax(s)
     tables(s)%cutrad = tcut(s)
     tables(s)%r = c_loc(r(toff(s)+1))
     tables(s)%f = c_loc(f(toff(s)+1))
     tables(s)%fp = c_loc(fp(toff(s)+1))
     tables(s)%fpp = c_loc(fpp(toff(s)+1))
  end do
  pm%nspecies = ns
  pm%tables = c_loc(tables(1))
  pm%nat = nat
  pm%xat = c_loc(xat(1))
  pm%species = c_loc(spc_t(1))

  npts = int(nodes%nstep(1),c_int64_t) * nodes%nstep(2) * nodes%nstep(3)
  allocate(crho(npts), cgrad(npts), rhoat(npts), rhofrag(max(npts*nfrag,1_c_int64_t)))
  ier = c2gpu_ctx_create(ctx, 0_c_int)
  if (ier /= C2GPU_OK) call die("c2gpu_ctx_create", ier)
  ier = c2gpu_nci_promol_nodes(ctx, nodes, x2c, pm, C2GPU_FP64, 0_c_int, c_loc(crho), c_loc(cgrad), c_loc(rhoat))
  if (ier /= C2GPU_OK) call die("c2gpu_nci_promol_nodes", ier)
  if (nfrag > 0) then
     ier = c2gpu_nci_fragments(ctx, nodes, ns, tables, nfrag, frag_nat, xyz, fspecies, C2GPU_FP64, 0_c_int, &
        c_loc(rhofrag))
     if (ier /= C2GPU_OK) call die("c2gpu_nci_fragments", ier)
  end if
  ! the cutoff pass over the node list as one axis, into a .dat buffer that
  ! holds a row (31 bytes) for every node
  n = (/ int(npts,c_int), 1_c_int, 1_c_int /)
  allocate(dat(31*npts))
  ier = c2gpu_nci_cutoff_frag(ctx, crho, cgrad, rhoat, rhofrag, nfrag, rhoparam, rhoparam2, n, C2GPU_FP64, &
     opts, 0_c_int, dat, 31*npts, dlen, sums)
  if (ier /= C2GPU_OK) call die("c2gpu_nci_cutoff_frag", ier)
  ier = c2gpu_ctx_destroy(ctx)

  open(newunit=lu, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(lu) dlen
  if (dlen > 0) write(lu) dat(1:dlen)
  write(lu) crho, cgrad, rhoat
  if (nfrag > 0) write(lu) rhofrag(1:npts*nfrag)
  write(lu) sums
  close(lu)

contains
  subroutine grow(a, m)
    real(c_double), allocatable, intent(inout) :: a(:)
    integer(c_int), intent(in) :: m
    real(c_double), allocatable :: t(:)
    allocate(t(size(a) + m))
    t(1:size(a)) = a
    call move_alloc(t, a)
  end subroutine grow

  subroutine die(what, code)
    character(*), intent(in) :: what
    integer(c_int), intent(in) :: code
    write(*,'("ERROR ",A," status ",I0)') what, code
    stop 1
  end subroutine die
end program c2gpu_nci_host
