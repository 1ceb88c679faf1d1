!------------------------------------------------------------------------------
! PROGRAM: dtc_gpu_driver
!------------------------------------------------------------------------------
!> Thin Fortran host for the gfx950 path: reads the reference's .meta keywords
!> (f-src_main/struct_process.f90:245-260) and the raw voxel image, forms the
!> domain decomposition (f-src/mod_decomp.f90:247-326, numbering of
!> f-src_main/struct_process.f90:901-910), solves the selected domains in one
!> batch through dtc_gpu and writes the 24 result leaves of
!> "Averaged Material Properties" (f-src_main/struct_process.f90:704-741) into
!> two typed streams, <basename>_results.int8.st and .real8.st, each domain at
!> slot lbound-1+(comm_nn-1)*size, with a text table of the leaves in
!> <basename>_results.head.
!>
!> Usage: dtc_gpu_driver <basename>.meta [precision 0|1]
!------------------------------------------------------------------------------
program dtc_gpu_driver

  use iso_c_binding
  use dtc_gpu
  implicit none

  character(len=512) :: meta, base, rawfile, line, key, typeraw, celt
  integer :: u, ios, i, j, k, nd, prec_arg, kind_b
  integer(c_int64_t) :: vdim(3), lo(3), up(3), x_D(3), nn_D(3), nn_i(3), xa(3), xe(3), llimit, dom
  real(c_double) :: spacing(3), sized(3), xdp(3), E, nu, strain
  integer :: macro_order
  integer(c_int) :: rc
  integer(1), allocatable, target :: img1(:,:,:)
  integer(2), allocatable, target :: img2(:,:,:)
  integer(4), allocatable, target :: img4(:,:,:)
  type(dtc_domain_desc), allocatable :: d(:)
  type(dtc_domain_result), allocatable :: res(:)
  type(dtc_solve_opts) :: o
  integer(c_int64_t), allocatable :: domains(:)
  ! per-domain voxel blocks kept alive until the solve returns
  type blk_t
     integer(1), allocatable :: b1(:,:,:)
     integer(2), allocatable :: b2(:,:,:)
     integer(4), allocatable :: b4(:,:,:)
  end type blk_t
  type(blk_t), allocatable, target :: blk(:)

  call get_command_argument(1, meta)
  if (len_trim(meta) == 0) stop "usage: dtc_gpu_driver <basename>.meta [precision]"
  prec_arg = 1
  call get_command_argument(2, line)
  if (len_trim(line) > 0) read(line, *) prec_arg
  base = meta(1:index(meta, ".meta", back=.true.) - 1)
  rawfile = trim(base)//".raw"

  ! ---- .meta keywords
  typeraw = "ik4"; celt = "HEX08"; macro_order = 1; strain = 1.e-6_c_double
  lo = 0; up = 0; llimit = 1
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
     end select
  end do
  close(u)

  ! ---- raw image, x fastest, signed kinds
  select case (trim(typeraw))
  case ("ik1"); kind_b = 1; allocate(img1(vdim(1), vdim(2), vdim(3)))
  case ("ik2"); kind_b = 2; allocate(img2(vdim(1), vdim(2), vdim(3)))
  case default; kind_b = 4; allocate(img4(vdim(1), vdim(2), vdim(3)))
  end select
  open(newunit=u, file=trim(rawfile), access="stream", form="unformatted", status="old", action="read")
  select case (kind_b)
  case (1); read(u) img1
  case (2); read(u) img2
  case default; read(u) img4
  end select
  close(u)

  ! ---- decomposition and domain list
  rc = dtc_ddc_params(sized, spacing, vdim, x_D, nn_D, xdp)
  if (rc /= DTC_OK) stop "ddc_params failed"
  nd = int((up(1)-lo(1)+1) * (up(2)-lo(2)+1) * (up(3)-lo(3)+1))
  allocate(domains(nd), d(nd), res(nd), blk(nd))
  nd = 0
  do k = int(lo(3)), int(up(3))
     do j = int(lo(2)), int(up(2))
        do i = int(lo(1)), int(up(1))
           nd = nd + 1
           domains(nd) = i + j * nn_D(1) + k * nn_D(1) * nn_D(2)
        end do
     end do
  end do

  do i = 1, nd
     dom = domains(i)
     rc = dtc_domain_bounds(dom, x_D, nn_D, nn_i, xa, xe)
     if (rc /= DTC_OK) stop "domain outside the decomposition"
     select case (kind_b)
     case (1)
        allocate(blk(i)%b1(x_D(1), x_D(2), x_D(3)))
        blk(i)%b1 = img1(xa(1):xe(1), xa(2):xe(2), xa(3):xe(3))
        d(i)%phi = c_loc(blk(i)%b1)
     case (2)
        allocate(blk(i)%b2(x_D(1), x_D(2), x_D(3)))
        blk(i)%b2 = img2(xa(1):xe(1), xa(2):xe(2), xa(3):xe(3))
        d(i)%phi = c_loc(blk(i)%b2)
     case default
        allocate(blk(i)%b4(x_D(1), x_D(2), x_D(3)))
        blk(i)%b4 = img4(xa(1):xe(1), xa(2):xe(2), xa(3):xe(3))
        d(i)%phi = c_loc(blk(i)%b4)
     end select
     d(i)%phi_kind = kind_b
     d(i)%micro_nen = merge(20, 8, trim(celt) == "HEX20")
     d(i)%macro_order = macro_order
     d(i)%n = x_D
     d(i)%llimit = llimit
     d(i)%xa_n = xa
     d(i)%delta = spacing
     d(i)%size_domain = sized(1)
     d(i)%E = E
     d(i)%nu = nu
     d(i)%rve_strain = strain
  end do

  o%precision = prec_arg
  rc = dtc_solve_batch(d, int(nd, c_int64_t), o, res)
  if (rc /= DTC_OK) then
     write(*, '(A,I0,2A)') "dtc_solve_batch failed: ", rc, " ", dtc_error_message()
     stop 1
  end if

  call write_results(trim(base)//"_results", nd, domains, res)
  do i = 1, nd
     write(*, '(A,I0,A,I0,A,I0,A,I0,A,ES10.3)') "domain ", domains(i), " status ", res(i)%status, &
          " elements ", res(i)%n_elems, " PCG iterations ", res(i)%iterations, " relres ", res(i)%relres_max
     write(*, '(6F12.4)') res(i)%C
  end do

contains

  !> 24 leaves of "Averaged Material Properties": int8 leaves 1-4, real8 leaves 5-24.
  !> Leaves 5/6 (start time, duration) and 9/13/15 (symmetry deviations of geb-lib's
  !> check_sym) and 10/11 (averaged stresses/strains, never computed by the
  !> reference) are written as zeros.
  subroutine write_results(stem, nd, domains, res)
    character(len=*), intent(in) :: stem
    integer, intent(in) :: nd
    integer(c_int64_t), intent(in) :: domains(:)
    type(dtc_domain_result), intent(in) :: res(:)
    integer, parameter :: NL = 24
    character(len=50), parameter :: desc(NL) = [character(len=50) :: &
         "Domain number", "Number of Elements", "Number of Nodes", "Collected logs", "Start Time", &
         "Duration", "Domain forces", "Effective numerical stiffness", &
         "Symmetry deviation - effective numerical stiffness", "Averaged stresses", "Averaged strains", &
         "Effective stiffness", "Symmetry deviation - effective stiffness", "Averaged Effective stiffness", &
         "Symmetry deviation - Averaged effective stiffness", "Rotation Angle CR_1", "Rotation Vector CR_1", &
         "Final coordinate system CR_1", "Optimized Effective stiffness CR_1", "Rotation Angle CR_2", &
         "Rotation Vector CR_2", "Final coordinate system CR_2", "Optimized Effective stiffness CR_2", &
         "Effective density"]
    integer :: sz(NL), nlc
    integer(c_int64_t) :: lb(NL), pos_i, pos_r
    integer :: ui, ur, uh, l, n
    integer(c_int64_t) :: ilog(24)
    real(c_double), allocatable :: zero(:)

    ! leaf sizes with no_lc = 24 (HEX08 macro element) or 60 (HEX20)
    nlc = merge(60, 24, macro_order == 2)
    sz = [1, 1, 1, 24, 1, 1, nlc*nlc, nlc*nlc, 1, 6*nlc, 6*nlc, 36, 1, 36, 1, 1, 3, 9, 36, 1, 3, 9, 36, 1]
    pos_i = 1; pos_r = 1
    do l = 1, NL
       if (l <= 4) then
          lb(l) = pos_i; pos_i = pos_i + int(sz(l), c_int64_t) * nd
       else
          lb(l) = pos_r; pos_r = pos_r + int(sz(l), c_int64_t) * nd
       end if
    end do
    open(newunit=ui, file=stem//".int8.st", access="stream", form="unformatted", status="replace")
    open(newunit=ur, file=stem//".real8.st", access="stream", form="unformatted", status="replace")
    allocate(zero(360)); zero = 0._c_double
    ilog = 0
    do n = 1, nd
       ! element offsets lbound-1+(comm_nn-1)*size, 8 bytes each
       write(ui, pos=8*(lb(1)-1+(n-1)*sz(1))+1) domains(n)
       write(ui, pos=8*(lb(2)-1+(n-1)*sz(2))+1) res(n)%n_elems
       write(ui, pos=8*(lb(3)-1+(n-1)*sz(3))+1) res(n)%n_nodes
       write(ui, pos=8*(lb(4)-1+(n-1)*sz(4))+1) ilog
       write(ur, pos=8*(lb(5)-1+(n-1)*sz(5))+1) zero(1:1)
       write(ur, pos=8*(lb(6)-1+(n-1)*sz(6))+1) zero(1:1)
       write(ur, pos=8*(lb(7)-1+(n-1)*sz(7))+1) res(n)%ff(1:sz(7))
       write(ur, pos=8*(lb(8)-1+(n-1)*sz(8))+1) res(n)%K_macro(1:sz(8))
       write(ur, pos=8*(lb(9)-1+(n-1)*sz(9))+1) zero(1:1)
       write(ur, pos=8*(lb(10)-1+(n-1)*sz(10))+1) zero(1:sz(10))
       write(ur, pos=8*(lb(11)-1+(n-1)*sz(11))+1) zero(1:sz(11))
       write(ur, pos=8*(lb(12)-1+(n-1)*sz(12))+1) res(n)%C
       write(ur, pos=8*(lb(13)-1+(n-1)*sz(13))+1) zero(1:1)
       write(ur, pos=8*(lb(14)-1+(n-1)*sz(14))+1) res(n)%EE
       write(ur, pos=8*(lb(15)-1+(n-1)*sz(15))+1) zero(1:1)
       write(ur, pos=8*(lb(16)-1+(n-1)*sz(16))+1) res(n)%rot_alpha(1)
       write(ur, pos=8*(lb(17)-1+(n-1)*sz(17))+1) res(n)%rot_n(:,1)
       write(ur, pos=8*(lb(18)-1+(n-1)*sz(18))+1) res(n)%rot_aa(:,:,1)
       write(ur, pos=8*(lb(19)-1+(n-1)*sz(19))+1) res(n)%rot_EE(:,:,1)
       write(ur, pos=8*(lb(20)-1+(n-1)*sz(20))+1) res(n)%rot_alpha(2)
       write(ur, pos=8*(lb(21)-1+(n-1)*sz(21))+1) res(n)%rot_n(:,2)
       write(ur, pos=8*(lb(22)-1+(n-1)*sz(22))+1) res(n)%rot_aa(:,:,2)
       write(ur, pos=8*(lb(23)-1+(n-1)*sz(23))+1) res(n)%rot_EE(:,:,2)
       write(ur, pos=8*(lb(24)-1+(n-1)*sz(24))+1) res(n)%density
    end do
    close(ui); close(ur)
    open(newunit=uh, file=stem//".head", status="replace", action="write")
    write(uh, '(A)') "Averaged Material Properties"
    do l = 1, NL
       write(uh, '(A50,1X,A5,3(1X,I12))') desc(l), merge("int8 ", "real8", l <= 4), int(sz(l)) * nd, &
            lb(l), lb(l) + int(sz(l), c_int64_t) * nd - 1
    end do
    close(uh)
  end subroutine write_results

end program dtc_gpu_driver
