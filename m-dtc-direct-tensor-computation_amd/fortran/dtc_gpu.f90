!------------------------------------------------------------------------------
! MODULE: dtc_gpu
!------------------------------------------------------------------------------
!> Fortran interface (iso_c_binding) of libdtc_amd.so, the gfx950 implementation
!> of the per-control-volume micro-FE path.  The derived types mirror the C
!> structs of include/dtc_capi.h field by field; ids are 1-based and arrays
!> column-major as in the reference.
!>
!> Drop-in points in the reference:
!>   dtc_mesh_count/dtc_mesh_build  <- call gen_quadmesh_from_phi, f-src/mod_struct_preprocess.f90:275-288
!>   dtc_boundaries                 <- call generate_boundaries,   f-src/mod_struct_preprocess.f90:424
!>   dtc_solve_domain/_batch        <- f-src/mod_dtc_main_subroutines.f90:426-1238 and
!>                                     f-src/mod_struct_calcmat.f90:252-1603
!>   dtc_ddc_params                 <- calc_general_ddc_params, f-src/mod_decomp.f90:247-326
!------------------------------------------------------------------------------
module dtc_gpu

  use iso_c_binding
  implicit none

  integer(c_int), parameter :: DTC_OK = 0, DTC_ERR_EMPTY_MESH = 1, DTC_ERR_FEW_CONSTRAINTS = 2, &
       DTC_ERR_OOM = 3, DTC_ERR_NO_CONVERGENCE = 4, DTC_ERR_BAD_ARGUMENT = 5, DTC_ERR_DEVICE = 6, &
       DTC_ERR_UNSUPPORTED = 7
  integer(c_int), parameter :: DTC_IK1 = 1, DTC_IK2 = 2, DTC_IK4 = 4
  integer(c_int), parameter :: DTC_FP64 = 0, DTC_FP32_REFINE = 1
  integer(c_int), parameter :: DTC_PC_JACOBI = 0, DTC_PC_MG = 1

  type, bind(c) :: dtc_domain_desc
     type(c_ptr)          :: phi = c_null_ptr
     integer(c_int)       :: phi_kind = DTC_IK4
     integer(c_int)       :: micro_nen = 8
     integer(c_int)       :: macro_order = 1
     integer(c_int)       :: pad = 0
     integer(c_int64_t)   :: n(3) = 0
     integer(c_int64_t)   :: llimit = 0
     integer(c_int64_t)   :: xa_n(3) = 1
     real(c_double)       :: delta(3) = 0._c_double
     real(c_double)       :: size_domain = 0._c_double
     real(c_double)       :: E = 0._c_double, nu = 0._c_double
     real(c_double)       :: rve_strain = 1.e-6_c_double
     type(c_ptr)          :: modulus = c_null_ptr
  end type dtc_domain_desc

  type, bind(c) :: dtc_solve_opts
     integer(c_int)       :: precision = DTC_FP32_REFINE
     integer(c_int)       :: rotation = 1
     real(c_double)       :: rel_tol = 1.e-8_c_double
     real(c_double)       :: inner_tol = 0._c_double    ! 0: automatic
     integer(c_int64_t)   :: max_iter = 200000
     real(c_double)       :: rot_step_deg = 1._c_double
     real(c_double)       :: num_zero = 1.e-9_c_double
     integer(c_int)       :: precond = DTC_PC_MG
     integer(c_int)       :: mg_degree = 0
  end type dtc_solve_opts

  type, bind(c) :: dtc_domain_result
     integer(c_int)       :: status, precision_used
     integer(c_int64_t)   :: n_elems, n_nodes, n_bnodes, iterations, outer_iterations
     real(c_double)       :: relres_max
     real(c_double)       :: ff(3600), K_macro(3600)   ! column-major R x R, R = 24 or 60
     real(c_double)       :: C(6,6), EE(6,6)
     real(c_double)       :: density
     real(c_double)       :: rot_alpha(2), rot_n(3,2), rot_aa(3,3,2), rot_EE(6,6,2)
     integer(c_int64_t)   :: rot_jj(2)
     real(c_double)       :: rot_crit(2)
  end type dtc_domain_result

  interface

     function dtc_last_error() bind(c, name="dtc_last_error") result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function dtc_last_error

     integer(c_int) function dtc_version() bind(c, name="dtc_version")
       import :: c_int
     end function dtc_version

     integer(c_int) function dtc_ddc_params(size_domain, delta, vdim, x_D, nn_D, x_D_phy) &
          bind(c, name="dtc_ddc_params")
       import :: c_int, c_int64_t, c_double
       real(c_double), intent(in) :: size_domain(3), delta(3)
       integer(c_int64_t), intent(in) :: vdim(3)
       integer(c_int64_t), intent(out) :: x_D(3), nn_D(3)
       real(c_double), intent(out) :: x_D_phy(3)
     end function dtc_ddc_params

     integer(c_int) function dtc_domain_bounds(domain, x_D, nn_D, nn_i, xa_n, xe_n) &
          bind(c, name="dtc_domain_bounds")
       import :: c_int, c_int64_t
       integer(c_int64_t), value :: domain
       integer(c_int64_t), intent(in) :: x_D(3), nn_D(3)
       integer(c_int64_t), intent(out) :: nn_i(3), xa_n(3), xe_n(3)
     end function dtc_domain_bounds

     integer(c_int) function dtc_extract_block(vol, kind, vdim, xa_n, xe_n, out) &
          bind(c, name="dtc_extract_block")
       import :: c_int, c_int64_t, c_ptr
       type(c_ptr), value :: vol, out
       integer(c_int), value :: kind
       integer(c_int64_t), intent(in) :: vdim(3), xa_n(3), xe_n(3)
     end function dtc_extract_block

     integer(c_int) function dtc_mesh_count(phi, kind, n, llimit, nen, n_elems, n_nodes) &
          bind(c, name="dtc_mesh_count")
       import :: c_int, c_int64_t, c_ptr
       type(c_ptr), value :: phi
       integer(c_int), value :: kind, nen
       integer(c_int64_t), intent(in) :: n(3)
       integer(c_int64_t), value :: llimit
       integer(c_int64_t), intent(out) :: n_elems, n_nodes
     end function dtc_mesh_count

     integer(c_int) function dtc_mesh_build(phi, kind, n, llimit, nen, delta, xa_n, elems_cap, nodes_cap, &
          nodes, elems, n_elems, n_nodes) bind(c, name="dtc_mesh_build")
       import :: c_int, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: phi
       integer(c_int), value :: kind, nen
       integer(c_int64_t), intent(in) :: n(3), xa_n(3)
       integer(c_int64_t), value :: llimit, elems_cap, nodes_cap
       real(c_double), intent(in) :: delta(3)
       real(c_double), intent(out) :: nodes(*)
       integer(c_int64_t), intent(out) :: elems(*)
       integer(c_int64_t), intent(out) :: n_elems, n_nodes
     end function dtc_mesh_build

     integer(c_int) function dtc_boundaries(nodes, n_nodes, min_c, max_c, dim_c, macro_order, n_b, ids, vals) &
          bind(c, name="dtc_boundaries")
       import :: c_int, c_int64_t, c_double, c_ptr
       real(c_double), intent(in) :: nodes(*), min_c(3), max_c(3), dim_c(3)
       integer(c_int64_t), value :: n_nodes
       integer(c_int), value :: macro_order
       integer(c_int64_t), intent(out) :: n_b
       type(c_ptr), value :: ids, vals
     end function dtc_boundaries

     integer(c_int) function dtc_element_stiffness(nen, E, nu, a, K) bind(c, name="dtc_element_stiffness")
       import :: c_int, c_double
       integer(c_int), value :: nen
       real(c_double), value :: E, nu, a
       real(c_double), intent(out) :: K(*)
     end function dtc_element_stiffness

     integer(c_int) function dtc_solve_domain(d, o, res, u, r) bind(c, name="dtc_solve_domain")
       import :: c_int, c_ptr, dtc_domain_desc, dtc_solve_opts, dtc_domain_result
       type(dtc_domain_desc), intent(in) :: d
       type(dtc_solve_opts), intent(in) :: o
       type(dtc_domain_result), intent(out) :: res
       type(c_ptr), value :: u, r
     end function dtc_solve_domain

     integer(c_int) function dtc_solve_batch(d, n_dom, o, res) bind(c, name="dtc_solve_batch")
       import :: c_int, c_int64_t, dtc_domain_desc, dtc_solve_opts, dtc_domain_result
       type(dtc_domain_desc), intent(in) :: d(*)
       integer(c_int64_t), value :: n_dom
       type(dtc_solve_opts), intent(in) :: o
       type(dtc_domain_result), intent(out) :: res(*)
     end function dtc_solve_batch

     integer(c_int) function dtc_rotation_search(EE, criterion, step_deg, num_zero, alpha, n, aa, EE_opt, jj, crit) &
          bind(c, name="dtc_rotation_search")
       import :: c_int, c_int64_t, c_double
       real(c_double), intent(in) :: EE(6,6)
       integer(c_int), value :: criterion
       real(c_double), value :: step_deg, num_zero
       real(c_double), intent(out) :: alpha, n(3), aa(3,3), EE_opt(6,6), crit
       integer(c_int64_t), intent(out) :: jj
     end function dtc_rotation_search

  end interface

contains

  !> dtc_last_error() as a Fortran string
  function dtc_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: cs(:)
    integer :: n
    type(c_ptr) :: p
    p = dtc_last_error()
    call c_f_pointer(p, cs, [1024])
    n = 0
    do while (n < 1024)
       if (cs(n+1) == c_null_char) exit
       n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(cs(1:n), msg)
  end function dtc_error_message

end module dtc_gpu
