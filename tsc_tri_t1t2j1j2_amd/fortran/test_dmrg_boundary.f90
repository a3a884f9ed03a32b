! Fortran host program over the drop-in subroutines (module tsc_dmrg): a
! cluster's blocks are grown from single sites with Get_Hamiltonian_Sys /
! _Env, the superblock [sys, s1 | env, s2] is solved with
! lanczos_memo_new, model_wave gives <psi|H psi>, and the truncation path
! (Get_density_matrix_sys, Get_truncation_operator keeping every state,
! Get_Hamiltonian_Sys) grows the sys block once more for a second
! superblock. Reads "N Ne" and then one bond "i j t J" per line from the
! file named on the command line; prints the numbers the test compares and,
! when a second file name is given, writes the untruncated sys and env
! blocks there (every sub-block of the Hamiltonian and of each carried
! site's c+, S, n) for a comparison with the reference's own numbers.
program test_dmrg_boundary
  use iso_c_binding
  use tsc_hotpath
  use tsc_dmrg
  implicit none
  integer :: nsite, ne, nb, k, ns, ios, u
  integer, allocatable :: bi(:), bj(:)
  real(8), allocatable :: bt(:), bjx(:)
  real(8) :: rq, nrm, dmax
  character(len=512) :: path
  type(Total_Model) :: sys, env, nxt
  type(Total_Basis) :: sbs, ebs
  type(Super_Basis) :: super
  type(Wavefunction) :: wv(1), hw
  type(Total_Block) :: trun, den, eff

  call get_command_argument(1, path)
  open(newunit=u, file=trim(path), status='old', action='read')
  read(u, *) nsite, ne, nb
  allocate(bi(nb), bj(nb), bt(nb), bjx(nb))
  do k = 1, nb
    read(u, *, iostat=ios) bi(k), bj(k), bt(k), bjx(k)
    if (ios /= 0) error stop 'bond list'
  end do
  close(u)
  call dmrg_setup_bonds(nsite, nb, bi, bj, bt, bjx)
  tot_num_up = ne
  tot_num_down = 0
  lanzero = 1.0d-13

  ! sys block: sites 1..N/2-1, env block: sites N..N/2+2, untruncated
  call Get_hamiltonian_site(1, sys)
  do k = 1, nsite / 2 - 2
    call Get_basis(sys%ham, sbs)
    call identity_trun(sbs, trun)
    call Get_Hamiltonian_Sys(sys, nxt, sbs, trun)
  end do
  call Get_hamiltonian_site(nsite, env)
  do k = 1, nsite / 2 - 2
    call Get_basis(env%ham, ebs)
    call identity_trun(ebs, trun)
    call Get_Hamiltonian_Env(env, nxt, ebs, trun)
  end do
  write(*, '(a,3i6)') 'LENS ', sys%len, env%len, sys%ham%num
  if (command_argument_count() >= 2) then
    call get_command_argument(2, path)
    open(newunit=u, file=trim(path), status='replace', action='write')
    call dump_model(u, 'sys', sys)
    call dump_model(u, 'env', env)
    close(u)
  end if

  ! update_trun_dia with every state kept returns the operator unchanged
  call Get_basis(sys%ham, sbs)
  call identity_trun(sbs, trun)
  call sector_trun(sys%ham, trun)
  call update_trun_dia(sys%ham, eff, trun)
  dmax = 0.0d0
  do k = 1, sys%ham%num
    dmax = max(dmax, maxval(abs(eff%sub(k)%mat - sys%ham%sub(k)%mat)))
  end do
  write(*, '(a,es12.4)') 'UOU_IDENTITY ', dmax

  ! superblock 1
  call Get_basis(sys%ham, sbs)
  call Get_basis(env%ham, ebs)
  call Get_super_basis(sbs, ebs, super, ne, 0)
  super%sys_len = sys%len + 1
  super%env_len = env%len + 1
  call allocate_wave(super, wv(1))
  call wave_initialize(wv(1), 7)
  call lanczos_memo_new(sys, env, sbs, ebs, super, wv, 1, 2)
  write(*, '(a,i10,f24.15)') 'E1 ', super%dim, Energys(1)
  call allocate_wave(super, hw)
  call model_wave(sys, env, sbs, ebs, wv(1), hw)
  rq = 0.0d0
  nrm = 0.0d0
  do k = 1, wv(1)%num
    rq = rq + real(sum(conjg(wv(1)%sub(k)%vec) * hw%sub(k)%vec))
    nrm = nrm + real(sum(conjg(wv(1)%sub(k)%vec) * wv(1)%sub(k)%vec))
  end do
  write(*, '(a,f24.15,es12.4)') 'RQ ', rq / nrm, abs(nrm - 1.0d0)

  ! truncation path: keep every state of rho_sys, grow sys by one site;
  ! env shrinks by one site (rebuilt from the end)
  kept_min = 100000
  kept_max = 100000
  call Get_density_matrix_sys(den, wv(1))
  call Get_truncation_operator(den, trun, 1.0d-14, ns)
  write(*, '(a,i8,f20.14)') 'KEPT ', ns, Entropys
  call Get_Hamiltonian_Sys(sys, nxt, sbs, trun)
  call Get_hamiltonian_site(nsite, env)
  do k = 1, nsite / 2 - 3
    call Get_basis(env%ham, ebs)
    call identity_trun(ebs, trun)
    call Get_Hamiltonian_Env(env, nxt, ebs, trun)
  end do
  call Get_basis(sys%ham, sbs)
  call Get_basis(env%ham, ebs)
  call Get_super_basis(sbs, ebs, super, ne, 0)
  super%sys_len = sys%len + 1
  super%env_len = env%len + 1
  call allocate_wave(super, wv(1))
  call wave_initialize(wv(1), 11)
  call lanczos_memo_new(sys, env, sbs, ebs, super, wv, 1, 2)
  write(*, '(a,i10,f24.15)') 'E2 ', super%dim, Energys(1)

contains

  ! one line per sub-block: tag, operator, num_up, num_down, down_dif,
  ! row_dim, sdim, then the matrix column-major as real, imaginary pairs
  subroutine dump_model(u, tag, m)
    integer, intent(in) :: u
    character(len=*), intent(in) :: tag
    type(Total_Model), intent(in) :: m
    integer :: k
    write(u, '(a,1x,i4,1x,i4)') 'MODEL '//tag, m%len, size(m%keep)
    write(u, '(a,1x,100i4)') 'KEEP '//tag, m%keep
    call dump_block(u, tag, 'H', 0, m%ham)
    do k = 1, size(m%keep)
      call dump_block(u, tag, 'c', m%keep(k), m%sub(k)%elec_up)
      call dump_block(u, tag, 'S', m%keep(k), m%sub(k)%spin_sd)
      call dump_block(u, tag, 'n', m%keep(k), m%sub(k)%num_sn)
    end do
  end subroutine dump_model

  subroutine dump_block(u, tag, name, site, b)
    integer, intent(in) :: u, site
    character(len=*), intent(in) :: tag, name
    type(Total_Block), intent(in) :: b
    integer :: q, i, j
    do q = 1, b%num
      write(u, '(a,1x,a,1x,i4,1x,i4,6(1x,i4))', advance='no') 'SUB '//tag, name, site, b%up_dif, &
        b%sub(q)%num_up, b%sub(q)%num_down, b%sub(q)%down_dif, b%sub(q)%row_dim, b%sub(q)%sdim
      do j = 1, b%sub(q)%sdim
        do i = 1, b%sub(q)%row_dim
          write(u, '(2(1x,es24.16))', advance='no') real(b%sub(q)%mat(i, j)), aimag(b%sub(q)%mat(i, j))
        end do
      end do
      write(u, *)
    end do
  end subroutine dump_block

  ! the identity truncation restricted to the block's own sectors
  subroutine sector_trun(ham, t)
    type(Total_Block), intent(in) :: ham
    type(Total_Block), intent(inout) :: t
    integer :: a, i
    call deallocate_block(t)
    allocate(t%sub(ham%num))
    t%num = ham%num
    do a = 1, ham%num
      t%sub(a)%num_up = ham%sub(a)%num_up
      t%sub(a)%num_down = ham%sub(a)%num_down
      t%sub(a)%row_dim = ham%sub(a)%sdim
      t%sub(a)%sdim = ham%sub(a)%sdim
      allocate(t%sub(a)%mat(ham%sub(a)%sdim, ham%sub(a)%sdim))
      t%sub(a)%mat = (0.0d0, 0.0d0)
      do i = 1, ham%sub(a)%sdim
        t%sub(a)%mat(i, i) = (1.0d0, 0.0d0)
      end do
    end do
  end subroutine sector_trun
end program test_dmrg_boundary
