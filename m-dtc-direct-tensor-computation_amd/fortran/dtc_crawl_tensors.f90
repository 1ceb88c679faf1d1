!------------------------------------------------------------------------------
! PROGRAM: dtc_crawl_tensors
!------------------------------------------------------------------------------
!> Collects the per-domain tensors that dtc_gpu_driver wrote into the puredat
!> tree <basename>_results.{head,int8.st,real8.st} into the tables the
!> reference's crawler writes (f-src_main/mrd_crawl_tensors.f90:16-580):
!>   <basename>.covo      effective stiffness      (leaf "Effective stiffness")
!>   <basename>.cr1       optimised tensor, CR_1    ("Optimized Effective stiffness CR_1",
!>                                                   position "Rotation Vector CR_1")
!>   <basename>.cr2       optimised tensor, CR_2    (likewise for CR_2)
!>   <basename>.covo.num  effective numerical stiffness, no_lc x no_lc
!> one row per computed domain of the .status tracker, in tracker order
!> (:200-215, :425-445).  Leaves are found by their descriptions in the .head.
!>
!> Row layout: the reference formats rows with geb-lib's
!> write_tensor_2nd_rank_R66_row / write_eff_num_stiff_row, which are absent, so
!> the column set here is this program's own (one header line names every
!> column), built from the fields the reference's crawler fills (:445-565):
!> domain number, section (domain position in the decomposition grid), its
!> physical bounds, domain size, density, elements, nodes, start time, duration,
!> symmetry deviation, micro element type, macro element order, optimisation
!> resolution (1, :428), the optimised position, the 24 collected logs and the
!> 36 tensor entries read as the reference reads them (:463-471, row index
!> outer).  The geb-lib measures doa_zener, doa_gebert and mps are not written.
!>
!> Usage: dtc_crawl_tensors <basename>.meta
!------------------------------------------------------------------------------
program dtc_crawl_tensors

  implicit none
  integer, parameter :: i8 = 8, r8 = 8, MCL = 512, MAXL = 64
  character(len=MCL) :: meta, base, line, key, celt
  integer(i8) :: lo(3), up(3), vdim(3), nd, grid(3), sec(3), stat_size
  real(r8) :: sized(3), spacing(3)
  integer :: u, ios, macro_order, nlc, nleaf, t, n, kk, ll, c, ucovo(3), unum
  character(len=64) :: ldesc(MAXL)
  integer :: ltype(MAXL)
  integer(i8) :: lno(MAXL), lb(MAXL)
  integer(i8), allocatable :: dstat(:), dom(:), nel(:), nno(:), logs(:)
  real(r8), allocatable :: dens(:), tst(:), tdu(:), ten(:), pos(:), knum(:), symv(:)
  integer :: ui, ur
  character(len=4), parameter :: suf(3) = ["covo", "cr1 ", "cr2 "]
  character(len=50), parameter :: tleaf(3) = [character(len=50) :: "Effective stiffness", &
       "Optimized Effective stiffness CR_1", "Optimized Effective stiffness CR_2"]
  character(len=50), parameter :: pleaf(3) = [character(len=50) :: "", "Rotation Vector CR_1", &
       "Rotation Vector CR_2"]
  character(len=50), parameter :: sleaf(3) = [character(len=50) :: "Symmetry deviation - effective stiffness", &
       "", ""]
  real(r8) :: mat(6, 6), p(3)
  integer(i8) :: slot, nrows
  logical :: ex

  call get_command_argument(1, meta)
  if (len_trim(meta) == 0) stop "usage: dtc_crawl_tensors <basename>.meta"
  base = meta(1:index(meta, ".meta", back=.true.) - 1)

  ! ---- .meta keywords (f-src_main/mrd_crawl_tensors.f90:115-125)
  macro_order = 1; celt = "HEX08"; lo = 0; up = 0
  open(newunit=u, file=trim(meta), status="old", action="read")
  do
     read(u, '(A)', iostat=ios) line
     if (ios /= 0) exit
     if (len_trim(line) < 3) cycle
     if (line(1:1) /= 'r' .and. line(1:1) /= 'w') cycle
     read(line(2:), *, iostat=ios) key
     if (ios /= 0) cycle
     select case (trim(key))
     case ("LO_BNDS_DMN_RANGE");  read(line(2:), *) key, lo
     case ("UP_BNDS_DMN_RANGE");  read(line(2:), *) key, up
     case ("MACRO_ELMNT_ORDER");  read(line(2:), *) key, macro_order
     case ("MICRO_ELMNT_TYPE");   read(line(2:), *) key, celt
     case ("SIZE_DOMAIN");        read(line(2:), *) key, sized
     case ("SPACING");            read(line(2:), *) key, spacing
     case ("DIMENSIONS");         read(line(2:), *) key, vdim
     end select
  end do
  close(u)
  if (macro_order == 1) then
     nlc = 24
  else if (macro_order == 2) then
     nlc = 60
  else
     stop "Macro element order not recognized. Choose '1' or '2'."
  end if
  nd = (up(1) - lo(1) + 1) * (up(2) - lo(2) + 1) * (up(3) - lo(3) + 1)
  ! domains per axis of the decomposition (x_D = NINT(SIZE/delta), bpoints = 1)
  grid = (vdim - 2) / nint(sized / spacing, i8)

  ! ---- status tracker (:190-215)
  inquire(file=trim(base)//".status", exist=ex, size=stat_size)
  if (.not. ex) stop "No status file found. Please check the input and/or computation."
  if (stat_size /= 8 * nd) stop "File size of the status file and the number of domains do not match."
  allocate(dstat(nd))
  open(newunit=u, file=trim(base)//".status", access="stream", form="unformatted", status="old", action="read")
  read(u) dstat
  close(u)

  ! ---- puredat head: every leaf's description, type, size and lower bound
  call read_head(trim(base)//"_results.head")
  open(newunit=ui, file=trim(base)//"_results.int8.st", access="stream", form="unformatted", status="old", &
       action="read")
  open(newunit=ur, file=trim(base)//"_results.real8.st", access="stream", form="unformatted", status="old", &
       action="read")
  call get_i("Domain number", dom)
  call get_i("Number of Elements", nel)
  call get_i("Number of Nodes", nno)
  call get_i("Collected logs", logs)
  call get_r("Start Time", tst)
  call get_r("Duration", tdu)
  call get_r("Effective density", dens)
  call get_r("Effective numerical stiffness", knum)

  do t = 1, 3
     open(newunit=ucovo(t), file=trim(base)//"."//trim(suf(t)), status="replace", action="write")
     call header(ucovo(t))
  end do
  open(newunit=unum, file=trim(base)//".covo.num", status="replace", action="write")
  write(unum, '(A)', advance="no") "Domain"
  do kk = 1, nlc
     do ll = 1, nlc
        write(unum, '(A,I0,A,I0)', advance="no") ",K", kk, "_", ll
     end do
  end do
  write(unum, '(A)') ""

  nrows = 0
  do t = 1, 3
     call get_r(tleaf(t), ten)
     if (t > 1) call get_r(pleaf(t), pos)
     if (t == 1) call get_r(sleaf(1), symv)
     ! rows in tracker order: the slot of each computed domain in the result leaves
     do n = 1, int(nd)
        if (dstat(n) < 0) cycle
        slot = -1
        do c = 1, size(dom)
           if (dom(c) == dstat(n) .and. nel(c) > 0) then
              slot = c
              exit
           end if
        end do
        if (slot < 0) cycle
        do kk = 1, 6
           do ll = 1, 6
              mat(kk, ll) = ten(36 * (slot - 1) + 6 * (kk - 1) + ll)
           end do
        end do
        p = 0._r8
        if (t > 1) p = pos(3 * (slot - 1) + 1:3 * slot)
        sec = [mod(dom(slot), grid(1)), mod(dom(slot) / grid(1), grid(2)), dom(slot) / (grid(1) * grid(2))]
        write(ucovo(t), '(I0,3(",",I0),8(",",ES24.16),2(",",I0),2(",",ES24.16))', advance="no") dom(slot), sec, &
             real(sec, r8) * sized(1), real(sec + 1, r8) * sized(1), sized(1), dens(slot), nel(slot), nno(slot), &
             tst(slot), tdu(slot)
        write(ucovo(t), '(",",ES24.16,",",A,",",I0,",",ES24.16,3(",",ES24.16))', advance="no") symdev(mat), &
             trim(celt), macro_order, 1._r8, p
        write(ucovo(t), '(24(",",I0))', advance="no") logs(24 * (slot - 1) + 1:24 * slot)
        write(ucovo(t), '(36(",",ES24.16))') ((mat(kk, ll), ll = 1, 6), kk = 1, 6)
        if (t == 1) then
           write(unum, '(I0)', advance="no") dom(slot)
           do c = 1, nlc * nlc
              write(unum, '(",",ES24.16)', advance="no") knum(int(nlc, i8) * nlc * (slot - 1) + c)
           end do
           write(unum, '(A)') ""
           nrows = nrows + 1
        end if
     end do
  end do
  do t = 1, 3
     close(ucovo(t))
  end do
  close(unum); close(ui); close(ur)
  write(*, '(A,I0,A)') "crawled ", nrows, " domains"

contains

  subroutine header(uh)
    integer, intent(in) :: uh
    integer :: a, b
    write(uh, '(A)', advance="no") "Domain,Section x,Section y,Section z,Lower x,Lower y,Lower z,Upper x,Upper y,"// &
         "Upper z,Domain size,Density,Elements,Nodes,Start time,Duration,Symmetry deviation,Micro element,"// &
         "Macro order,Opt. resolution,Position 1,Position 2,Position 3"
    do a = 1, 24
       write(uh, '(A,I0)', advance="no") ",Log ", a
    end do
    do a = 1, 6
       do b = 1, 6
          write(uh, '(A,I0,I0)', advance="no") ",C", a, b
       end do
    end do
    write(uh, '(A)') ""
  end subroutine header

  real(r8) function symdev(a)
    real(r8), intent(in) :: a(6, 6)
    real(r8) :: s
    s = maxval(abs(a))
    symdev = 0._r8
    if (s > 0._r8) symdev = maxval(abs(a - transpose(a))) / s
  end function symdev

  subroutine read_head(path)
    character(len=*), intent(in) :: path
    integer :: uh, q1, q2
    logical :: inleaf
    character(len=MCL) :: l
    nleaf = 0
    inleaf = .false.
    open(newunit=uh, file=path, status="old", action="read")
    do
       read(uh, '(A)', iostat=ios) l
       if (ios /= 0) exit
       if (index(l, "<--leaf-->") == 1) then
          nleaf = nleaf + 1
          if (nleaf > MAXL) stop "too many leaves in the head"
          inleaf = .true.
       else if (index(l, "<==branch==>") == 1) then
          inleaf = .false.
       else if (inleaf .and. index(l, "<description>") == 1) then
          q1 = index(l, "'")
          q2 = index(l, "'", back=.true.)
          ldesc(nleaf) = l(q1 + 1:q2 - 1)
       else if (inleaf .and. index(l, "<no_of_data>") == 1) then
          read(l(13:), *) lno(nleaf)
       else if (inleaf .and. index(l, "<type_of_data>") == 1) then
          read(l(15:), *) ltype(nleaf)
       else if (inleaf .and. index(l, "<lower_bound>") == 1) then
          read(l(14:), *) lb(nleaf)
       end if
    end do
    close(uh)
  end subroutine read_head

  integer function find(desc)
    character(len=*), intent(in) :: desc
    integer :: l
    do l = 1, nleaf
       if (trim(ldesc(l)) == trim(desc)) then
          find = l
          return
       end if
    end do
    write(*, '(3A)') "leaf '", trim(desc), "' not in the head"
    stop 1
  end function find

  subroutine get_i(desc, v)
    character(len=*), intent(in) :: desc
    integer(i8), allocatable, intent(out) :: v(:)
    integer :: l
    l = find(desc)
    if (ltype(l) /= 4) stop "leaf is not int8"
    allocate(v(lno(l)))
    read(ui, pos=8 * (lb(l) - 1) + 1) v
  end subroutine get_i

  subroutine get_r(desc, v)
    character(len=*), intent(in) :: desc
    real(r8), allocatable, intent(out) :: v(:)
    integer :: l
    l = find(desc)
    if (ltype(l) /= 5) stop "leaf is not real8"
    allocate(v(lno(l)))
    read(ur, pos=8 * (lb(l) - 1) + 1) v
  end subroutine get_r

end program dtc_crawl_tensors
