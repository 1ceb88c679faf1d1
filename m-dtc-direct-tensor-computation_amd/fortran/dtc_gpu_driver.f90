!------------------------------------------------------------------------------
! PROGRAM: dtc_gpu_driver
!------------------------------------------------------------------------------
!> Thin Fortran host for the gfx950 path, in place of the master/worker program
!> f-src_main/struct_process.f90:
!>  * reads the .meta keywords (f-src_main/struct_process.f90:245-260);
!>  * forms the domain decomposition (f-src/mod_decomp.f90:247-326) and the
!>    domain numbers of the selected range (f-src_main/struct_process.f90:901-910);
!>  * keeps the restart tracker <basename>.status, one int8 per domain, -999
!>    until the domain is computed, then its domain number; an existing tracker
!>    of another size is an error, domains already computed are skipped
!>    (f-src_main/struct_process.f90:358-367,481-535,1266-1268,1357);
!>  * streams each domain's voxel block from the raw image (only the block is
!>    read, f-src/mod_struct_preprocess.f90:165-251) and solves the domains in
!>    groups of DTC_DRIVER_GROUP (default 256) through one dtc_solve_batch per
!>    group, so the image is never held in memory whole;
!>  * writes the puredat tree of the reference's output (f-src/mod_puredat.f90:
!>    4630-4700 write_tree/write_branch/write_leaf, stream names :2467-2480):
!>    <stem>.head with the branches "Input parameters" (18 leaves,
!>    f-src_main/struct_process.f90:393-414), "Global domain decomposition"
!>    (8 leaves, f-src/mod_decomp.f90:276-318) and "Averaged Material
!>    Properties" (24 leaves, f-src_main/struct_process.f90:700-741), and the
!>    typed streams <stem>.int8.st, .real8.st and .char.st.  A domain's result
!>    goes to element lbound-1+(n-1)*size of its leaf, n its position in the
!>    selected range (one "communicator": domains_per_comm = 2 x domains, the
!>    reference's over-allocation, :350-352).
!>
!> Usage: dtc_gpu_driver <basename>.meta [precision 0|1]
!>   reads <basename>.raw, writes <basename>_results.{head,int8.st,real8.st,char.st}
!>   and <basename>.status.
!------------------------------------------------------------------------------
program dtc_gpu_driver

  use iso_c_binding
  use dtc_gpu
  implicit none

  integer, parameter :: MCL = 512, NLEAF = 18 + 8 + 24
  character(len=MCL) :: meta, base, rawfile, line, key, typeraw, stem, outfmt, outamount, restart, statfile
  character(len=8) :: celt
  integer :: u, ios, i, j, k, nd, prec_arg, kind_b, macro_order, parts, g0, g1, ng, group, nlc
  integer(c_int64_t) :: vdim(3), lo(3), up(3), x_D(3), nn_D(3), nn_i(3), xa(3), xe(3), llimit
  integer(c_int64_t) :: bpoints(3), dpc, stat_size
  real(c_double) :: spacing(3), sized(3), xdp(3), E, nu, strain
  integer(c_int) :: rc
  type(dtc_domain_desc), allocatable :: d(:)
  type(dtc_domain_result), allocatable :: res(:)
  type(dtc_solve_opts) :: o
  integer(c_int64_t), allocatable :: domains(:), dstat(:)
  integer, allocatable :: todo(:)
  logical :: stat_exists, head_exists, fresh
  ! per-domain voxel blocks of one group, kept alive until its solve returns
  type blk_t
     integer(1), allocatable :: b1(:,:,:)
     integer(2), allocatable :: b2(:,:,:)
     integer(4), allocatable :: b4(:,:,:)
  end type blk_t
  type(blk_t), allocatable, target :: blk(:)
  ! the puredat tree: leaf descriptions, types (4 int8, 5 real8, 6 char), sizes, bounds
  character(len=50) :: ldesc(NLEAF)
  integer :: ltype(NLEAF), lsize(NLEAF)
  integer(c_int64_t) :: lno(NLEAF), lb(NLEAF), st_size(7)
  integer :: ui, ur, uc
  integer(c_int64_t) :: clk0, clk1, clk_rate, t_unix0, t_unix1

  call get_command_argument(1, meta)
  if (len_trim(meta) == 0) stop "usage: dtc_gpu_driver <basename>.meta [precision]"
  prec_arg = 1
  call get_command_argument(2, line)
  if (len_trim(line) > 0) read(line, *) prec_arg
  base = meta(1:index(meta, ".meta", back=.true.) - 1)
  rawfile = trim(base)//".raw"
  stem = trim(base)//"_results"
  statfile = trim(base)//".status"
  group = 256
  call get_environment_variable("DTC_DRIVER_GROUP", line)
  if (len_trim(line) > 0) read(line, *) group
  group = max(group, 1)

  ! ---- .meta keywords
  typeraw = "ik4"; celt = "HEX08"; macro_order = 1; strain = 1.e-6_c_double
  lo = 0; up = 0; llimit = 1; parts = 1; outfmt = "FMPS"; outamount = "PRODUCTION"; restart = "N"
  open(newunit=u, file=trim(meta), status="old", action="read")
  do
     read(u, '(A)', iostat=ios) line
     if (ios /= 0) exit
     if (len_trim(line) < 3) cycle
     if (line(1:1) /= 'r' .and. line(1:1) /= 'w') cycle
     read(line(2:), *, iostat=ios) key
     if (ios /= 0) cycle
     select case (trim(key))
     case ("TYPE_RAW");           read(line(2:), *) key, typeraw
     case ("DIMENSIONS");         read(line(2:), *) key, vdim
     case ("SPACING");            read(line(2:), *) key, spacing
     case ("SIZE_DOMAIN");        read(line(2:), *) key, sized
     case ("LO_BNDS_DMN_RANGE");  read(line(2:), *) key, lo
     case ("UP_BNDS_DMN_RANGE");  read(line(2:), *) key, up
     case ("BINARIZE_LO");        read(line(2:), *) key, llimit
     case ("RVE_STRAIN");         read(line(2:), *) key, strain
     case ("YOUNG_MODULUS");      read(line(2:), *) key, E
     case ("POISSON_RATIO");      read(line(2:), *) key, nu
     case ("MACRO_ELMNT_ORDER");  read(line(2:), *) key, macro_order
     case ("MICRO_ELMNT_TYPE");   read(line(2:), *) key, celt
     case ("MESH_PER_SUB_DMN");   read(line(2:), *) key, parts
     case ("OUT_FMT");            read(line(2:), *) key, outfmt
     case ("RESTART");            read(line(2:), *) key, restart
     end select
  end do
  close(u)
  select case (trim(typeraw))
  case ("ik1"); kind_b = 1
  case ("ik2"); kind_b = 2
  case default; kind_b = 4
  end select
  nlc = merge(60, 24, macro_order == 2)

  ! ---- decomposition and domain list (x fastest)
  rc = dtc_ddc_params(sized, spacing, vdim, x_D, nn_D, xdp)
  if (rc /= DTC_OK) stop "ddc_params failed"
  bpoints = 1
  nd = int((up(1)-lo(1)+1) * (up(2)-lo(2)+1) * (up(3)-lo(3)+1))
  allocate(domains(nd), dstat(nd))
  nd = 0
  do k = int(lo(3)), int(up(3))
     do j = int(lo(2)), int(up(2))
        do i = int(lo(1)), int(up(1))
           nd = nd + 1
           domains(nd) = i + j * nn_D(1) + k * nn_D(1) * nn_D(2)
        end do
     end do
  end do
  dpc = 2_c_int64_t * nd

  ! ---- restart tracker
  inquire(file=trim(statfile), exist=stat_exists, size=stat_size)
  if (stat_exists .and. stat_size /= 8_c_int64_t * nd) then
     write(*, '(A)') "Size of the status file and the number of requested domains do not match: "// &
          "the domain range changed since the tracker was written; set up a new job."
     stop 2
  end if
  inquire(file=trim(stem)//".head", exist=head_exists)
  fresh = .not. (stat_exists .and. head_exists)
  if (fresh) then
     dstat = -999_c_int64_t
  else
     open(newunit=u, file=trim(statfile), access="stream", form="unformatted", status="old", action="read")
     read(u) dstat
     close(u)
  end if

  ! ---- puredat tree and streams (a new job writes them whole, a restart reopens them)
  call build_tree()
  if (fresh) then
     call write_head(trim(stem)//".head")
     open(newunit=ui, file=trim(stem)//".int8.st", access="stream", form="unformatted", status="replace")
     open(newunit=ur, file=trim(stem)//".real8.st", access="stream", form="unformatted", status="replace")
     open(newunit=uc, file=trim(stem)//".char.st", access="stream", form="unformatted", status="replace")
     call write_parameters()
     call zero_results()
     open(newunit=u, file=trim(statfile), access="stream", form="unformatted", status="replace")
     write(u) dstat
     close(u)
  else
     open(newunit=ui, file=trim(stem)//".int8.st", access="stream", form="unformatted", status="old")
     open(newunit=ur, file=trim(stem)//".real8.st", access="stream", form="unformatted", status="old")
     open(newunit=uc, file=trim(stem)//".char.st", access="stream", form="unformatted", status="old")
  end if

  ! ---- domains still to compute, in groups
  allocate(todo(nd))
  ng = 0
  do i = 1, nd
     if (dstat(i) < 0) then
        ng = ng + 1
        todo(ng) = i
     end if
  end do
  write(*, '(A,I0,A,I0,A)') "domains: ", nd, " selected, ", nd - ng, " already computed"
  allocate(d(group), res(group), blk(group))
  o%precision = prec_arg
  call system_clock(count_rate=clk_rate)
  do g0 = 1, ng, group
     g1 = min(ng, g0 + group - 1)
     do i = g0, g1
        call load_domain(todo(i), i - g0 + 1)
     end do
     call system_clock(clk0)
     t_unix0 = time()
     rc = dtc_solve_batch(d, int(g1 - g0 + 1, c_int64_t), o, res)
     if (rc /= DTC_OK) then
        write(*, '(A,I0,2A)') "dtc_solve_batch failed: ", rc, " ", dtc_error_message()
        stop 1
     end if
     call system_clock(clk1)
     t_unix1 = time()
     do i = g0, g1
        call write_domain(todo(i), res(i - g0 + 1), real(clk0, c_double) / real(clk_rate, c_double), &
             real(clk1 - clk0, c_double) / real(clk_rate, c_double) / real(g1 - g0 + 1, c_double))
        if (res(i - g0 + 1)%status == DTC_OK) dstat(todo(i)) = domains(todo(i))
        write(*, '(A,I0,A,I0,A,I0,A,I0,A,ES10.3)') "domain ", domains(todo(i)), " status ", &
             res(i - g0 + 1)%status, " elements ", res(i - g0 + 1)%n_elems, " PCG iterations ", &
             res(i - g0 + 1)%iterations, " relres ", res(i - g0 + 1)%relres_max
        write(*, '(6F12.4)') res(i - g0 + 1)%C
     end do
     flush(ui); flush(ur)
     ! the tracker follows each group, so an interrupted job restarts after it
     open(newunit=u, file=trim(statfile), access="stream", form="unformatted", status="old")
     write(u, pos=1) dstat
     close(u)
     do i = 1, g1 - g0 + 1
        if (allocated(blk(i)%b1)) deallocate(blk(i)%b1)
        if (allocated(blk(i)%b2)) deallocate(blk(i)%b2)
        if (allocated(blk(i)%b4)) deallocate(blk(i)%b4)
     end do
  end do
  close(ui); close(ur); close(uc)

contains

  !> Reads the voxel block of selected domain n into slot s (row by row from the
  !> raw stream: x_D(1) voxels per (y, z) row), and fills its descriptor.
  subroutine load_domain(n, s)
    integer, intent(in) :: n, s
    integer(c_int64_t) :: y, z, off
    integer :: uraw
    rc = dtc_domain_bounds(domains(n), x_D, nn_D, nn_i, xa, xe)
    if (rc /= DTC_OK) stop "domain outside the decomposition"
    open(newunit=uraw, file=trim(rawfile), access="stream", form="unformatted", status="old", action="read")
    select case (kind_b)
    case (1); allocate(blk(s)%b1(x_D(1), x_D(2), x_D(3)))
    case (2); allocate(blk(s)%b2(x_D(1), x_D(2), x_D(3)))
    case default; allocate(blk(s)%b4(x_D(1), x_D(2), x_D(3)))
    end select
    do z = xa(3), xe(3)
       do y = xa(2), xe(2)
          off = ((z - 1) * vdim(2) * vdim(1) + (y - 1) * vdim(1) + xa(1) - 1) * kind_b + 1
          select case (kind_b)
          case (1); read(uraw, pos=off) blk(s)%b1(:, y - xa(2) + 1, z - xa(3) + 1)
          case (2); read(uraw, pos=off) blk(s)%b2(:, y - xa(2) + 1, z - xa(3) + 1)
          case default; read(uraw, pos=off) blk(s)%b4(:, y - xa(2) + 1, z - xa(3) + 1)
          end select
       end do
    end do
    close(uraw)
    select case (kind_b)
    case (1); d(s)%phi = c_loc(blk(s)%b1)
    case (2); d(s)%phi = c_loc(blk(s)%b2)
    case default; d(s)%phi = c_loc(blk(s)%b4)
    end select
    d(s)%phi_kind = kind_b
    d(s)%micro_nen = merge(20, 8, trim(celt) == "HEX20")
    d(s)%macro_order = macro_order
    d(s)%n = x_D
    d(s)%llimit = llimit
    d(s)%xa_n = xa
    d(s)%delta = spacing
    d(s)%size_domain = sized(1)
    d(s)%E = E
    d(s)%nu = nu
    d(s)%rve_strain = strain
  end subroutine load_domain

  !> Leaves of the tree in write order, their types and element counts; bounds are
  !> assigned per stream in that order (set_bounds_in_branch, f-src/mod_puredat.f90:2007).
  subroutine build_tree()
    integer :: l, t
    character(len=50), parameter :: dres(24) = [character(len=50) :: &
         "Domain number", "Number of Elements", "Number of Nodes", "Collected logs", "Start Time", &
         "Duration", "Domain forces", "Effective numerical stiffness", &
         "Symmetry deviation - effective numerical stiffness", "Averaged stresses", "Averaged strains", &
         "Effective stiffness", "Symmetry deviation - effective stiffness", "Averaged Effective stiffness", &
         "Symmetry deviation - Averaged effective stiffness", "Rotation Angle CR_1", "Rotation Vector CR_1", &
         "Final coordinate system CR_1", "Optimized Effective stiffness CR_1", "Rotation Angle CR_2", &
         "Rotation Vector CR_2", "Final coordinate system CR_2", "Optimized Effective stiffness CR_2", &
         "Effective density"]
    ldesc(1:18) = [character(len=50) :: "muCT puredat pro_path", "muCT puredat pro_name", &
         "Physical domain size", "Lower bounds of selected domain range", "Upper bounds of selected domain range", &
         "Grid spacings", "Lower limit of iso value", "Element type  on micro scale", &
         "No of mesh parts per subdomain", "Output Format", "Average strain on RVE", "Young_s modulus", &
         "Poisson_s ratio", "Element order on macro scale", "Output amount", "Restart", &
         "Number of voxels per direction", "Domains per communicator"]
    ltype(1:18) = [6, 6, 5, 4, 4, 5, 4, 6, 4, 6, 5, 5, 5, 4, 6, 6, 4, 4]
    lno(1:18) = [integer(c_int64_t) :: MCL, MCL, 3, 3, 3, 3, 1, 8, 1, 8, 1, 1, 1, 1, MCL, 1, 3, 1]
    ldesc(19:26) = [character(len=50) :: "x_D", "x_D_phy", "delta", "nn_D", "nn_D_max", "x_VD", "no_dat_D", &
         "bpoints"]
    ltype(19:26) = [4, 5, 5, 4, 4, 4, 4, 4]
    lno(19:26) = [integer(c_int64_t) :: 3, 3, 3, 3, 1, 3, 1, 3]
    ldesc(27:50) = dres
    lsize(27:50) = [1, 1, 1, 24, 1, 1, nlc*nlc, nlc*nlc, 1, 6*nlc, 6*nlc, 36, 1, 36, 1, 1, 3, 9, 36, 1, 3, 9, 36, 1]
    ltype(27:50) = [4, 4, 4, 4, (5, l = 5, 24)]
    lno(27:50) = int(lsize(27:50), c_int64_t) * dpc
    lsize(1:26) = 0
    st_size = 0
    do l = 1, NLEAF
       t = ltype(l)
       lb(l) = st_size(t) + 1
       st_size(t) = st_size(t) + lno(l)
    end do
  end subroutine build_tree

  subroutine write_head(path)
    character(len=*), intent(in) :: path
    integer :: uh, b, l, l0, l1
    character(len=40), parameter :: bname(3) = [character(len=40) :: "Input parameters", &
         "Global domain decomposition", "Averaged Material Properties"]
    open(newunit=uh, file=path, status="replace", action="write")
    write(uh, '(A)') "<==branch==>"
    write(uh, '(4(A))') "<description> ", "'", "results", "'"
    write(uh, '(A,I0)') "<no_of_branches> ", 3
    write(uh, '(A,I0)') "<no_of_leaves> ", 0
    write(uh, '(A,L1)') "<streams_allocated> ", .true.
    write(uh, '(A,I0)') "<size_int1_stream> ", st_size(1)
    write(uh, '(A,I0)') "<size_int2_stream> ", st_size(2)
    write(uh, '(A,I0)') "<size_int4_stream> ", st_size(3)
    write(uh, '(A,I0)') "<size_int8_stream> ", st_size(4)
    write(uh, '(A,I0)') "<size_real_stream> ", st_size(5)
    write(uh, '(A,I0)') "<size_char_stream> ", st_size(6)
    write(uh, '(A,I0)') "<size_log_stream>  ", st_size(7)
    do b = 1, 3
       l0 = merge(1, merge(19, 27, b == 2), b == 1)
       l1 = merge(18, merge(26, 50, b == 2), b == 1)
       write(uh, '(A)') "<==branch==>"
       write(uh, '(4(A))') "<description> ", "'", trim(bname(b)), "'"
       write(uh, '(A,I0)') "<no_of_branches> ", 0
       write(uh, '(A,I0)') "<no_of_leaves> ", l1 - l0 + 1
       write(uh, '(A,L1)') "<streams_allocated> ", .false.
       do l = l0, l1
          write(uh, '(A)') "<--leaf-->"
          write(uh, '(4(A))') "<description> ", "'", trim(ldesc(l)), "'"
          write(uh, '(A,I0)') "<no_of_data> ", lno(l)
          write(uh, '(A,I0)') "<type_of_data> ", ltype(l)
          write(uh, '(A,I0)') "<lower_bound> ", lb(l)
          write(uh, '(A,I0)') "<upper_bound> ", lb(l) + lno(l) - 1
       end do
    end do
    close(uh)
  end subroutine write_head

  subroutine put_i(l, off, v)
    integer, intent(in) :: l
    integer(c_int64_t), intent(in) :: off, v(:)
    write(ui, pos=8*(lb(l)-1+off)+1) v
  end subroutine put_i

  subroutine put_r(l, off, v)
    integer, intent(in) :: l
    integer(c_int64_t), intent(in) :: off
    real(c_double), intent(in) :: v(:)
    write(ur, pos=8*(lb(l)-1+off)+1) v
  end subroutine put_r

  subroutine put_c(l, v)
    integer, intent(in) :: l
    character(len=*), intent(in) :: v
    character(len=MCL) :: s
    s = v
    write(uc, pos=lb(l)) s(1:lno(l))
  end subroutine put_c

  !> leaves of "Input parameters" and "Global domain decomposition"
  subroutine write_parameters()
    character(len=MCL) :: ppath, pname
    integer :: sl
    sl = index(base, "/", back=.true.)
    ppath = base(1:sl)
    pname = base(sl+1:)
    call put_c(1, ppath)
    call put_c(2, pname)
    call put_r(3, 0_c_int64_t, sized)
    call put_i(4, 0_c_int64_t, lo)
    call put_i(5, 0_c_int64_t, up)
    call put_r(6, 0_c_int64_t, spacing)
    call put_i(7, 0_c_int64_t, [llimit])
    call put_c(8, celt)
    call put_i(9, 0_c_int64_t, [int(parts, c_int64_t)])
    call put_c(10, outfmt)
    call put_r(11, 0_c_int64_t, [strain])
    call put_r(12, 0_c_int64_t, [E])
    call put_r(13, 0_c_int64_t, [nu])
    call put_i(14, 0_c_int64_t, [int(macro_order, c_int64_t)])
    call put_c(15, outamount)
    call put_c(16, restart(1:1))
    call put_i(17, 0_c_int64_t, vdim)
    call put_i(18, 0_c_int64_t, [dpc])
    call put_i(19, 0_c_int64_t, x_D)
    call put_r(20, 0_c_int64_t, xdp)
    call put_r(21, 0_c_int64_t, spacing)
    call put_i(22, 0_c_int64_t, nn_D)
    call put_i(23, 0_c_int64_t, [nn_D(1) * nn_D(2) * nn_D(3) - 1])
    call put_i(24, 0_c_int64_t, nn_D * x_D)
    call put_i(25, 0_c_int64_t, [(x_D(1) + 2*bpoints(1)) * (x_D(2) + 2*bpoints(2)) * (x_D(3) + 2*bpoints(3))])
    call put_i(26, 0_c_int64_t, bpoints)
  end subroutine write_parameters

  !> the result leaves of a new job start as zeros over their whole size
  subroutine zero_results()
    integer :: l
    integer(c_int64_t) :: zi(1024)
    real(c_double) :: zr(1024)
    integer(c_int64_t) :: m, c
    zi = 0; zr = 0._c_double
    do l = 27, NLEAF
       m = 0
       do while (m < lno(l))
          c = min(1024_c_int64_t, lno(l) - m)
          if (ltype(l) == 4) then
             call put_i(l, m, zi(1:c))
          else
             call put_r(l, m, zr(1:c))
          end if
          m = m + c
       end do
    end do
  end subroutine zero_results

  !> max |A - A^T| / max |A|: the symmetry deviation written to leaves 9, 13 and
  !> 15.  The reference's check_sym (geb-lib) is absent; this measure is
  !> documented as its substitute (parity unpinned).
  real(c_double) function symdev(a, m)
    integer, intent(in) :: m
    real(c_double), intent(in) :: a(m, m)
    real(c_double) :: s
    s = maxval(abs(a))
    symdev = 0._c_double
    if (s > 0._c_double) symdev = maxval(abs(a - transpose(a))) / s
  end function symdev

  !> the 24 result leaves of selected domain n (slot n-1 of each leaf)
  subroutine write_domain(n, r, t_start, t_dur)
    integer, intent(in) :: n
    type(dtc_domain_result), intent(in) :: r
    real(c_double), intent(in) :: t_start, t_dur
    integer(c_int64_t) :: s, logs(24)
    real(c_double) :: km(nlc, nlc)
    integer :: l
    s = n - 1
    ! Collected logs (f-src/mod_struct_calcmat.f90:181-183,1657-1659): time stamps of the
    ! batch start and end, stage flags 15..21 (1 = solved), one GPU, rank 0, their sum
    logs = 0
    logs(1) = t_unix0
    logs(7) = t_unix1
    if (r%status == DTC_OK) logs(15:21) = 1
    logs(22) = 1
    logs(23) = 0
    logs(24) = sum(logs(15:21))
    call put_i(27, s * lsize(27), [domains(n)])
    call put_i(28, s * lsize(28), [r%n_elems])
    call put_i(29, s * lsize(29), [r%n_nodes])
    call put_i(30, s * lsize(30), logs)
    call put_r(31, s * lsize(31), [t_start])
    call put_r(32, s * lsize(32), [t_dur])
    call put_r(33, s * lsize(33), r%ff(1:nlc*nlc))
    call put_r(34, s * lsize(34), r%K_macro(1:nlc*nlc))
    km = reshape(r%K_macro(1:nlc*nlc), [nlc, nlc])
    call put_r(35, s * lsize(35), [symdev(km, nlc)])
    call put_r(38, s * lsize(38), reshape(r%C, [36]))
    call put_r(39, s * lsize(39), [symdev(r%C, 6)])
    call put_r(40, s * lsize(40), reshape(r%EE, [36]))
    call put_r(41, s * lsize(41), [symdev(r%EE, 6)])
    call put_r(42, s * lsize(42), [r%rot_alpha(1)])
    call put_r(43, s * lsize(43), r%rot_n(:, 1))
    call put_r(44, s * lsize(44), reshape(r%rot_aa(:, :, 1), [9]))
    call put_r(45, s * lsize(45), reshape(r%rot_EE(:, :, 1), [36]))
    call put_r(46, s * lsize(46), [r%rot_alpha(2)])
    call put_r(47, s * lsize(47), r%rot_n(:, 2))
    call put_r(48, s * lsize(48), reshape(r%rot_aa(:, :, 2), [9]))
    call put_r(49, s * lsize(49), reshape(r%rot_EE(:, :, 2), [36]))
    call put_r(50, s * lsize(50), [r%density])
  end subroutine write_domain

end program dtc_gpu_driver
