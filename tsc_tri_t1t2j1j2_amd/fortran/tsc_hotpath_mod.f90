!=============================================================================
! Fortran binding of the MI355X hot path (include/tsc_hotpath.h).
!
! A Fortran host (the reference's DMRG_code) calls the HIP library through
! these iso_c_binding interfaces. Derived types tsc_contrib / tsc_outblock
! mirror the C structs byte for byte; device objects are opaque c_ptr
! handles. Status codes: 0 = ok; tsc_check turns anything else into
! `error stop` with the library's message (the reference stops on
! inconsistent quantum numbers the same way, get_sys.f90:3166-3168).
!=============================================================================
module tsc_hotpath
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: TSC_PREC_F64 = 0, TSC_PREC_F32 = 1, TSC_PREC_BF16X3 = 2
  integer(c_int8_t), parameter :: TSC_OP_N = 0, TSC_OP_T = 1, TSC_OP_C = 2, TSC_OP_CONJ = 3, TSC_OP_I = 4
  integer(c_int8_t), parameter :: TSC_BUF_OPS = 0, TSC_BUF_IN = 1, TSC_BUF_MID = 2, TSC_BUF_OUT = 3

  type, bind(c) :: tsc_contrib
    integer(c_int64_t) :: a_off, b_off
    integer(c_int32_t) :: lda, ldb
    integer(c_int32_t) :: k
    integer(c_int8_t)  :: a_buf, b_buf, a_op, b_op
    real(c_double)     :: coef_re, coef_im
  end type tsc_contrib

  type, bind(c) :: tsc_outblock
    integer(c_int64_t) :: c_off
    integer(c_int32_t) :: ldc, m, n
    integer(c_int32_t) :: c_buf, cbeg, cend, mode
  end type tsc_outblock

  interface
    function tsc_last_error() bind(c, name="tsc_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    integer(c_int) function tsc_ctx_create(device, ctx) bind(c, name="tsc_ctx_create")
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr) :: ctx
    end function
    integer(c_int) function tsc_ctx_destroy(ctx) bind(c, name="tsc_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tsc_vec_create(ctx, n, v) bind(c, name="tsc_vec_create")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      type(c_ptr) :: v
    end function
    integer(c_int) function tsc_vec_destroy(v) bind(c, name="tsc_vec_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: v
    end function
    integer(c_int) function tsc_vec_upload(v, host) bind(c, name="tsc_vec_upload")
      import :: c_int, c_ptr
      type(c_ptr), value :: v, host
    end function
    integer(c_int) function tsc_vec_download(v, host) bind(c, name="tsc_vec_download")
      import :: c_int, c_ptr
      type(c_ptr), value :: v, host
    end function
    integer(c_int) function tsc_vec_dot_re(u, v, r) bind(c, name="tsc_vec_dot_re")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: u, v
      real(c_double) :: r
    end function
    integer(c_int) function tsc_plan_create(ctx, ops, ops_elems, mid_elems, c1, nc1, o1, no1, c2, nc2, o2, no2, &
                                            in_elems, out_elems, plan) bind(c, name="tsc_plan_create")
      import :: c_int, c_ptr, c_int64_t, c_int32_t
      type(c_ptr), value :: ctx, ops, c1, o1, c2, o2
      integer(c_int64_t), value :: ops_elems, mid_elems, in_elems, out_elems
      integer(c_int32_t), value :: nc1, no1, nc2, no2
      type(c_ptr) :: plan
    end function
    integer(c_int) function tsc_plan_destroy(plan) bind(c, name="tsc_plan_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
    end function
    integer(c_int) function tsc_hpsi(plan, vin, vout, prec) bind(c, name="tsc_hpsi")
      import :: c_int, c_ptr
      type(c_ptr), value :: plan, vin, vout
      integer(c_int), value :: prec
    end function
    integer(c_int) function tsc_lanczos(plan, psi, tol, max_inner, max_outer, prec, refine_iters, energy, info) &
        bind(c, name="tsc_lanczos")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: plan, psi
      real(c_double), value :: tol
      integer(c_int), value :: max_inner, max_outer, prec, refine_iters
      real(c_double) :: energy
      integer(c_int32_t) :: info(2)
    end function
    integer(c_int) function tsc_rho_truncate(ctx, psi, nsec, sys_dim, env_dim, twos, side, kept_min, kept_max, &
                                             trun_err, eigval, kept, u, scalars) bind(c, name="tsc_rho_truncate")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx, psi, sys_dim, env_dim, twos, eigval, kept, u, scalars
      integer(c_int32_t), value :: nsec, kept_min, kept_max
      integer(c_int), value :: side
      real(c_double), value :: trun_err
    end function
    integer(c_int) function tsc_update_trun(ctx, nblk, row, col, kl, kr, ori_off, ul_off, ur_off, eff_off, &
                                            ori, ori_elems, u, u_elems, eff, eff_elems, prec) &
        bind(c, name="tsc_update_trun")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx, row, col, kl, kr, ori_off, ul_off, ur_off, eff_off, ori, u, eff
      integer(c_int32_t), value :: nblk
      integer(c_int64_t), value :: ori_elems, u_elems, eff_elems
      integer(c_int), value :: prec
    end function
    real(c_double) function tsc_w6j(j1, j2, j3, l1, l2, l3) bind(c, name="tsc_w6j")
      import :: c_double, c_int32_t
      integer(c_int32_t), value :: j1, j2, j3, l1, l2, l3
    end function
  end interface

contains

  subroutine tsc_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(tsc_last_error(), msg, [1024])
    n = 0
    do while (n < 1024)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write(*, *) 'tsc error in ', where, ': ', msg(1:n)
    error stop 'tsc hot path call failed'
  end subroutine tsc_check

  ! update_trun_dia/_ndia for one sub-block: eff = U_lhs^H . O . U_rhs
  ! (block.f90:3511-3514, 3672-3675) on the device.
  subroutine tsc_uou_block(ctx, ori, ul, ur, eff)
    type(c_ptr), intent(in) :: ctx
    complex(c_double_complex), intent(in), target :: ori(:, :), ul(:, :), ur(:, :)
    complex(c_double_complex), intent(out), target :: eff(:, :)
    integer(c_int32_t), target :: row(1), col(1), kl(1), kr(1)
    integer(c_int64_t), target :: oo(1), lo(1), ro(1), eo(1)
    complex(c_double_complex), allocatable, target :: ubuf(:)
    row(1) = size(ori, 1); col(1) = size(ori, 2); kl(1) = size(ul, 2); kr(1) = size(ur, 2)
    oo(1) = 0; lo(1) = 0; ro(1) = size(ul); eo(1) = 0
    allocate(ubuf(size(ul) + size(ur)))
    ubuf(1:size(ul)) = reshape(ul, [size(ul)])
    ubuf(size(ul) + 1:) = reshape(ur, [size(ur)])
    call tsc_check(tsc_update_trun(ctx, 1_c_int32_t, c_loc(row), c_loc(col), c_loc(kl), c_loc(kr), c_loc(oo), &
                                   c_loc(lo), c_loc(ro), c_loc(eo), c_loc(ori), int(size(ori), c_int64_t), &
                                   c_loc(ubuf), int(size(ubuf), c_int64_t), c_loc(eff), int(size(eff), c_int64_t), &
                                   TSC_PREC_F64), 'update_trun')
    deallocate(ubuf)
  end subroutine tsc_uou_block

end module tsc_hotpath
