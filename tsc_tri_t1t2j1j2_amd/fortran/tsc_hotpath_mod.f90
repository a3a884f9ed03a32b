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

  ! descriptors of the SU(2) planner and of block models (C structs of the
  ! same names in tsc_hotpath.h, byte for byte)
  type, bind(c) :: tsc_sector
    integer(c_int32_t) :: n, twos, dim
  end type tsc_sector

  type, bind(c) :: tsc_op_block
    integer(c_int32_t) :: bra, ket
    integer(c_int64_t) :: off
  end type tsc_op_block

  type, bind(c) :: tsc_op_desc
    integer(c_int32_t) :: side, up_dif, k2, nblk
    type(c_ptr) :: blk
  end type tsc_op_desc

  type, bind(c) :: tsc_pair_desc
    real(c_double) :: coef
    integer(c_int32_t) :: block_op, block_tilde, site_op, site_tilde
  end type tsc_pair_desc

  type, bind(c) :: tsc_term_desc
    real(c_double) :: coef
    integer(c_int32_t) :: l_kind, l_op, l_tilde, r_kind, r_op, r_tilde
  end type tsc_term_desc

  type, bind(c) :: tsc_sb_desc
    integer(c_int32_t) :: nsys, nenv
    type(c_ptr) :: sys, env
    integer(c_int64_t) :: arena_len
    integer(c_int32_t) :: nops
    type(c_ptr) :: ops, sys_ham, env_ham
    integer(c_int32_t) :: nlpairs
    type(c_ptr) :: lpairs
    integer(c_int32_t) :: nrpairs
    type(c_ptr) :: rpairs
    integer(c_int32_t) :: nterms
    type(c_ptr) :: terms
    integer(c_int32_t) :: ntot, s2tot
  end type tsc_sb_desc

  type, bind(c) :: tsc_site_ops
    integer(c_int32_t) :: site
    type(c_ptr) :: op_c, op_s, op_n
  end type tsc_site_ops

  type, bind(c) :: tsc_model_desc
    integer(c_int32_t) :: nsec
    type(c_ptr) :: sec
    integer(c_int32_t) :: nsites
    type(c_ptr) :: sites, ham
    integer(c_int32_t) :: nkeep
    type(c_ptr) :: keep
  end type tsc_model_desc

  type, bind(c) :: tsc_bond
    integer(c_int32_t) :: i, j
    real(c_double) :: t, jex
  end type tsc_bond

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
                                             trun_err, eigval, kept, u, u_cap, scalars) bind(c, name="tsc_rho_truncate")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx, psi, sys_dim, env_dim, twos, eigval, kept, u, scalars
      integer(c_int32_t), value :: nsec, kept_min, kept_max
      integer(c_int), value :: side
      real(c_double), value :: trun_err
      integer(c_int64_t), value :: u_cap
    end function
    integer(c_int) function tsc_select_states(nsec, dim, twos, eigval, kept_min, kept_max, trun_err, kept, &
                                              scalars) bind(c, name="tsc_select_states")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: dim, twos, eigval, kept, scalars
      integer(c_int32_t), value :: nsec, kept_min, kept_max
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
    integer(c_int) function tsc_version() bind(c, name="tsc_version")
      import :: c_int
    end function
    integer(c_int) function tsc_ctx_sync(ctx) bind(c, name="tsc_ctx_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tsc_ctx_mfma_layout(ctx, modes3) bind(c, name="tsc_ctx_mfma_layout")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: modes3(3)
    end function
    integer(c_int) function tsc_ctx_mem_info(ctx, free_bytes, total_bytes) bind(c, name="tsc_ctx_mem_info")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t) :: free_bytes, total_bytes
    end function
    integer(c_int) function tsc_vec_norm(v, r) bind(c, name="tsc_vec_norm")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: v
      real(c_double) :: r
    end function
    function tsc_vec_dptr(v) bind(c, name="tsc_vec_dptr") result(p)
      import :: c_ptr
      type(c_ptr), value :: v
      type(c_ptr) :: p
    end function
    integer(c_int) function tsc_vec_fill_gaussian(v, seed, key, sigma) bind(c, name="tsc_vec_fill_gaussian")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: v
      integer(c_int64_t), value :: seed, key
      real(c_double), value :: sigma
    end function
    integer(c_int) function tsc_vec_axpby(y, a, x, b) bind(c, name="tsc_vec_axpby")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: y, x
      real(c_double), value :: a, b
    end function
    integer(c_int) function tsc_vec_copy(y, x) bind(c, name="tsc_vec_copy")
      import :: c_int, c_ptr
      type(c_ptr), value :: y, x
    end function
    integer(c_int) function tsc_vec_copy_range(y, y_off, x, x_off, n) bind(c, name="tsc_vec_copy_range")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: y, x
      integer(c_int64_t), value :: y_off, x_off, n
    end function
    integer(c_int) function tsc_vec_upload_range(v, off, host, n) bind(c, name="tsc_vec_upload_range")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: v, host
      integer(c_int64_t), value :: off, n
    end function
    integer(c_int) function tsc_vec_download_range(v, off, host, n) bind(c, name="tsc_vec_download_range")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: v, host
      integer(c_int64_t), value :: off, n
    end function
    integer(c_int) function tsc_plan_create_dev(ctx, ops, mid_elems, c1, nc1, o1, no1, c2, nc2, o2, no2, &
                                                in_elems, out_elems, plan) bind(c, name="tsc_plan_create_dev")
      import :: c_int, c_ptr, c_int64_t, c_int32_t
      type(c_ptr), value :: ctx, ops, c1, o1, c2, o2
      integer(c_int64_t), value :: mid_elems, in_elems, out_elems
      integer(c_int32_t), value :: nc1, no1, nc2, no2
      type(c_ptr) :: plan
    end function
    integer(c_int) function tsc_plan_update_ops(plan, ops) bind(c, name="tsc_plan_update_ops")
      import :: c_int, c_ptr
      type(c_ptr), value :: plan, ops
    end function
    integer(c_int) function tsc_plan_fill_ops_gaussian(plan, off, count, seed, key, sigma) &
        bind(c, name="tsc_plan_fill_ops_gaussian")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: plan
      integer(c_int64_t), value :: off, count, seed, key
      real(c_double), value :: sigma
    end function
    integer(c_int) function tsc_plan_flops(plan, flops) bind(c, name="tsc_plan_flops")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double) :: flops
    end function
    integer(c_int) function tsc_hpsi_host(plan, vin, vout, prec) bind(c, name="tsc_hpsi_host")
      import :: c_int, c_ptr
      type(c_ptr), value :: plan, vin, vout
      integer(c_int), value :: prec
    end function
    integer(c_int) function tsc_hpsi_loop(plan, vin, vout) bind(c, name="tsc_hpsi_loop")
      import :: c_int, c_ptr
      type(c_ptr), value :: plan, vin, vout
    end function
    ! in_c128 / out_c128: device addresses (tsc_vec_dptr); allocates nothing
    integer(c_int) function tsc_hpsi_noalloc(plan, in_c128, out_c128) bind(c, name="tsc_hpsi_noalloc")
      import :: c_int, c_ptr
      type(c_ptr), value :: plan, in_c128, out_c128
    end function
    integer(c_int) function tsc_hpsi_graph(plan, vin, vout, prec) bind(c, name="tsc_hpsi_graph")
      import :: c_int, c_ptr
      type(c_ptr), value :: plan, vin, vout
      integer(c_int), value :: prec
    end function
    integer(c_int) function tsc_plan_last_times(plan, ms1, ms2) bind(c, name="tsc_plan_last_times")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: plan
      real(c_float) :: ms1, ms2
    end function
    integer(c_int) function tsc_rho_last_times(ctx, ms4) bind(c, name="tsc_rho_last_times")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      real(c_float) :: ms4(4)
    end function
    integer(c_int) function tsc_plan_autotune(plan, vin, vout, prec, flags, chosen) bind(c, name="tsc_plan_autotune")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan, vin, vout
      integer(c_int), value :: prec, flags
      real(c_double) :: chosen
    end function
    integer(c_int) function tsc_lanczos_replays(plan, replays) bind(c, name="tsc_lanczos_replays")
      import :: c_int, c_ptr
      type(c_ptr), value :: plan, replays
    end function
    integer(c_int) function tsc_lanczos_refinement(plan, rec5) bind(c, name="tsc_lanczos_refinement")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double) :: rec5(5)
    end function
    integer(c_int) function tsc_density_matrix(ctx, psi, nsec, sys_dim, env_dim, side, rho) &
        bind(c, name="tsc_density_matrix")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx, psi, sys_dim, env_dim, rho
      integer(c_int32_t), value :: nsec
      integer(c_int), value :: side
    end function
    integer(c_int) function tsc_truncation_operator(ctx, nsec, dim, twos, rho, kept_min, kept_max, trun_err, &
                                                    eigval, kept, u, u_cap, scalars) &
                                                    bind(c, name="tsc_truncation_operator")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx, dim, twos, rho, eigval, kept, u, scalars
      integer(c_int32_t), value :: nsec, kept_min, kept_max
      real(c_double), value :: trun_err
      integer(c_int64_t), value :: u_cap
    end function
    integer(c_int) function tsc_tasks_flops(tasks, flops2) bind(c, name="tsc_tasks_flops")
      import :: c_int, c_ptr
      type(c_ptr), value :: tasks, flops2
    end function
    integer(c_int) function tsc_model_tasks(sys, env, s1, s2, nbonds, bonds, ntot, s2tot, arena, arena_len, &
                                            tasks) bind(c, name="tsc_model_tasks")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: sys, env, bonds, arena
      integer(c_int32_t), value :: s1, s2, nbonds, ntot, s2tot
      integer(c_int64_t), value :: arena_len
      type(c_ptr) :: tasks
    end function
    integer(c_int) function tsc_renorm_tasks(blk, new_site, nbonds, bonds, nkeep, keep_sites, norder, order, kept, &
                                             arena, arena_len, tasks) bind(c, name="tsc_renorm_tasks")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: blk, bonds, keep_sites, order, kept, arena
      integer(c_int32_t), value :: new_site, nbonds, nkeep, norder
      integer(c_int64_t), value :: arena_len
      type(c_ptr) :: tasks
    end function
    integer(c_int) function tsc_tasks_layout(tasks, counts3, layout, offsets, new_sec, enl_sec) &
                                             bind(c, name="tsc_tasks_layout")
      import :: c_int, c_ptr
      type(c_ptr), value :: tasks, counts3, layout, offsets, new_sec, enl_sec
    end function
    integer(c_int) function tsc_tasks_lengths(tasks, lens3) bind(c, name="tsc_tasks_lengths")
      import :: c_int, c_ptr
      type(c_ptr), value :: tasks, lens3
    end function
    integer(c_int) function tsc_tasks_copy_arena(tasks, arena) bind(c, name="tsc_tasks_copy_arena")
      import :: c_int, c_ptr
      type(c_ptr), value :: tasks, arena
    end function
    integer(c_int) function tsc_superblock_tasks(desc, tasks) bind(c, name="tsc_superblock_tasks")
      import :: c_int, c_ptr
      type(c_ptr), value :: desc
      type(c_ptr) :: tasks
    end function
    integer(c_int) function tsc_tasks_info(tasks, counts9) bind(c, name="tsc_tasks_info")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: tasks
      integer(c_int64_t) :: counts9(9)
    end function
    integer(c_int) function tsc_tasks_copy(tasks, c1, o1, c2, o2, super_xy, lsec, rsec) bind(c, name="tsc_tasks_copy")
      import :: c_int, c_ptr
      type(c_ptr), value :: tasks, c1, o1, c2, o2, super_xy, lsec, rsec
    end function
    integer(c_int) function tsc_tasks_plan(ctx, tasks, arena, arena_len, plan) bind(c, name="tsc_tasks_plan")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, tasks, arena
      integer(c_int64_t), value :: arena_len
      type(c_ptr) :: plan
    end function
    integer(c_int) function tsc_tasks_plan_dev(ctx, tasks, arena, plan) bind(c, name="tsc_tasks_plan_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, tasks, arena
      type(c_ptr) :: plan
    end function
    integer(c_int) function tsc_tasks_destroy(tasks) bind(c, name="tsc_tasks_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: tasks
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
