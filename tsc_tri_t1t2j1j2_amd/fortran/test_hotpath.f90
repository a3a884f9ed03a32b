! Fortran host check of the hot path through iso_c_binding: a Hermitian
! block operator psi -> A psi + psi A^T solved by the device Lanczos, a
! U^H O U renormalisation, a 6j symbol, and rho / eigensolve / kept-state
! selection of a two-sector psi (Get_density_matrix_sys +
! Get_truncation_operator). Prints PASS/FAIL lines.
program test_hotpath
  use iso_c_binding
  use tsc_hotpath
  implicit none
  integer, parameter :: n = 12
  type(c_ptr) :: ctx, plan, psi
  complex(c_double_complex), target :: a(n, n), x(n * n), ori(5, 4), ul(5, 3), ur(4, 2), eff(3, 2), ref(3, 2)
  type(tsc_contrib), target :: c(2)
  type(tsc_outblock), target :: o(1)
  real(c_double) :: e, th
  integer(c_int32_t) :: info(2)
  integer :: i
  real(c_double) :: re(n, n)
  complex(c_double_complex), target :: w(10), u(10)
  integer(c_int32_t), target :: sdim(2), edim(2), twos(2), kept(2)
  real(c_double), target :: eig(5), scal(4)
  type(c_ptr) :: wv

  ! A = D T D^H with T the tridiagonal (2, -1) Toeplitz matrix and D a
  ! diagonal phase matrix: Hermitian, lowest eigenvalue 2 - 2 cos(pi/(n+1))
  a = (0d0, 0d0)
  do i = 1, n
    a(i, i) = 2d0
    if (i < n) then
      a(i, i + 1) = -exp(cmplx(0d0, 0.3d0 * i - 0.3d0 * (i + 1), c_double_complex))
      a(i + 1, i) = conjg(a(i, i + 1))
    end if
  end do
  call tsc_check(tsc_ctx_create(0_c_int, ctx), 'ctx_create')
  c(1) = tsc_contrib(0_c_int64_t, 0_c_int64_t, n, n, n, TSC_BUF_OPS, TSC_BUF_IN, TSC_OP_N, TSC_OP_N, 1d0, 0d0)
  c(2) = tsc_contrib(0_c_int64_t, 0_c_int64_t, n, n, n, TSC_BUF_IN, TSC_BUF_OPS, TSC_OP_N, TSC_OP_T, 1d0, 0d0)
  o(1) = tsc_outblock(0_c_int64_t, n, n, n, 3, 0, 2, 0)
  call tsc_check(tsc_plan_create(ctx, c_loc(a), int(n * n, c_int64_t), 0_c_int64_t, c_null_ptr, 0_c_int32_t, &
                                 c_null_ptr, 0_c_int32_t, c_loc(c), 2_c_int32_t, c_loc(o), 1_c_int32_t, &
                                 int(n * n, c_int64_t), int(n * n, c_int64_t), plan), 'plan_create')
  do i = 1, n * n
    x(i) = cmplx(cos(0.1d0 * i), sin(0.37d0 * i), c_double_complex)
  end do
  call tsc_check(tsc_vec_create(ctx, int(n * n, c_int64_t), psi), 'vec_create')
  call tsc_check(tsc_vec_upload(psi, c_loc(x)), 'vec_upload')
  call tsc_check(tsc_lanczos(plan, psi, 1d-12, 100, 4, TSC_PREC_F64, 0, e, info), 'lanczos')
  ! exact: twice the lowest eigenvalue of A
  th = 2d0 - 2d0 * cos(acos(-1d0) / (n + 1))
  if (abs(e - 2d0 * th) < 1d-9) then
    write(*, '(a,2f20.12)') 'PASS lanczos ', e, 2d0 * th
  else
    write(*, '(a,2f20.12)') 'FAIL lanczos ', e, 2d0 * th
  end if
  ! U^H O U
  call random_number(re(1:5, 1:4)); ori = cmplx(re(1:5, 1:4), 0.3d0, c_double_complex)
  call random_number(re(1:5, 1:3)); ul = cmplx(re(1:5, 1:3), -0.2d0, c_double_complex)
  call random_number(re(1:4, 1:2)); ur = cmplx(re(1:4, 1:2), 0.1d0, c_double_complex)
  call tsc_uou_block(ctx, ori, ul, ur, eff)
  ref = matmul(conjg(transpose(ul)), matmul(ori, ur))
  if (maxval(abs(eff - ref)) < 1d-12) then
    write(*, '(a)') 'PASS update_trun'
  else
    write(*, '(a,es12.3)') 'FAIL update_trun ', maxval(abs(eff - ref))
  end if
  ! {1/2 1/2 1; 1/2 1/2 0} = 1/2 ... doubled arguments
  th = tsc_w6j(1, 1, 2, 1, 1, 0)
  if (abs(th - 0.5d0) < 1d-14) then
    write(*, '(a)') 'PASS w6j'
  else
    write(*, '(a,f12.8)') 'FAIL w6j ', th
  end if
  ! psi sectors (sys 3 x env 2) and (2 x 2) with singular values^2 0.5, 0.1 and
  ! 0.3, 0.1; kept_min 1, kept_max 2 (the kept rule adds states past kept_min
  ! while the discarded weight exceeds trun_err, den_mat.f90:301-316): the two
  ! largest states, one per sector, discarded weight 0.2
  w = (0d0, 0d0)
  w(1) = sqrt(0.5d0); w(5) = sqrt(0.1d0)
  w(7) = sqrt(0.3d0); w(10) = sqrt(0.1d0)
  sdim = [3, 2]; edim = [2, 2]; twos = [0, 0]
  call tsc_check(tsc_vec_create(ctx, 10_c_int64_t, wv), 'vec_create')
  call tsc_check(tsc_vec_upload(wv, c_loc(w)), 'vec_upload')
  call tsc_check(tsc_rho_truncate(ctx, wv, 2_c_int32_t, c_loc(sdim), c_loc(edim), c_loc(twos), 0_c_int, &
                                  1_c_int32_t, 2_c_int32_t, 1d-12, c_loc(eig), c_loc(kept), c_loc(u), 10_c_int64_t, &
                                  c_loc(scal)), &
                 'rho_truncate')
  if (all(kept == [1, 1]) .and. nint(scal(1)) == 2 .and. abs(scal(2) - 0.2d0) < 1d-12 .and. &
      abs(eig(3) - 0.5d0) < 1d-12 .and. abs(eig(5) - 0.3d0) < 1d-12 .and. abs(abs(u(1)) - 1d0) < 1d-12) then
    write(*, '(a)') 'PASS rho_truncate'
  else
    write(*, '(a,2i4,5f10.5)') 'FAIL rho_truncate ', kept, scal(1), scal(2), eig(3), eig(5), abs(u(1))
  end if
  call tsc_check(tsc_vec_destroy(wv), 'vec_destroy')
  call tsc_check(tsc_vec_destroy(psi), 'vec_destroy')
  call tsc_check(tsc_plan_destroy(plan), 'plan_destroy')
  call tsc_check(tsc_ctx_destroy(ctx), 'ctx_destroy')
end program test_hotpath
