!=============================================================================
! Drop-in subroutines of the DMRG hot path with the reference's names and
! argument lists (hyx15903/TSC_Tri_t1t2J1J2, DMRG_code/), over derived types
! that carry the same data as the reference's (common.f90:84-174):
!
!   model_wave(sys,env,sys_bs,env_bs,invec,outvec)        get_sys.f90:6915
!   hmerge(sys,env,sys_bs,env_bs)                          get_sys.f90:7398
!   lanczos_memo_new(sys,env,sys_bs,env_bs,super,eigvec,levels,num_idx)
!                                                          eigen.f90:4
!   Get_density_matrix_sys(sys_den,wave) / _env(env_den,wave)
!                                                          den_mat.f90:4,55
!   Get_truncation_operator(den_mat,trun_op,trun_err,num_state)
!                                                          den_mat.f90:184
!   update_trun_dia(ori,eff,trun) / update_trun_ndia(ori,eff,trun)
!                                                          block.f90:3465,3535
!   Get_Hamiltonian_Sys(pre,next,basis,trun) / _Env       get_sys.f90:4,305
!
! plus the bookkeeping they rest on (Get_hamiltonian_site, Get_basis,
! Get_super_basis, allocate_wave, wave_initialize). Every numerical stage
! runs in the HIP library through the C-ABI (module tsc_hotpath): the term
! list, hmerge's merged partner operators and the contraction plans are
! built by tsc_model_tasks / tsc_renorm_tasks, H.psi and Lanczos run on the
! device, rho and its eigensolve too.
!
! Conventions. Quantum numbers (num_up, num_down) = (N, 2S); the site basis
! is electron (1,1) then hole (0,0). A block's sectors are the sub-blocks
! of its Hamiltonian, in order (every sector has a diagonal block). An
! operator's sub(k) is keyed by its ket sector (num_up, num_down) and
! down_dif; the bra sector is (num_up + up_dif, num_down + down_dif);
! mat(row_dim = bra dim, sdim = ket dim). The types hold the reference's
! numbers, env side included: env-block operators as the reference stores
! them (it applies them as psi . O^H, get_sys.f90:1869), env truncations as
! its rho_env eigenvectors, and wave functions whose enlarged env basis puts
! the site's fermion modes before the block's, as its
! update_block_ndia_env / update_site_ndia_env signs imply
! (block.f90:2828-3459). The library applies env operators as psi . O^T and
! orders an enlarged env block then site, so the boundary converts: env
! operators are conjugated when packed and when stored, env truncation
! matrices are conjugated and their rows, like the env columns of a wave
! function, multiplied by (-1)^(N_block N_site) of their enlarged-env
! product state (env_signs). The model carries, for every site with a bond
! leaving the block, that site's
! c+ (elec_up), S (spin_sd) and n (num_sn); `sites` lists the block's sites
! and `keep` the carried ones. Bonds (i, j, t, J) are set once with
! dmrg_setup_bonds (the role of dmrg_setup's neighbour tables,
! dmrg_sub.f90:453-540).
!=============================================================================
module tsc_dmrg
  use iso_c_binding
  use tsc_hotpath
  implicit none

  type Sub_Block
    integer :: num_up = 0, num_down = 0, down_dif = 0, row_dim = 0, sdim = 0
    complex(c_double_complex), allocatable :: mat(:, :)
  end type Sub_Block

  type Total_Block
    integer :: len = 0, num = 0, dim = 0, up_dif = 0, down_dif = 0
    type(Sub_Block), allocatable :: sub(:)
  end type Total_Block

  type Block_Set
    type(Total_Block) :: elec_up, spin_sd, num_sn
  end type Block_Set

  type Total_Model
    integer :: len = 0
    type(Total_Block) :: ham
    integer, allocatable :: sites(:)
    integer, allocatable :: keep(:)
    type(Block_Set), allocatable :: sub(:)
  end type Total_Model

  type Sub_Basis_Map
    integer :: spos = 0, sdim = 0, bl_num_up = 0, bl_num_down = 0, st_num_up = 0, st_num_down = 0
  end type Sub_Basis_Map

  type Basis_Map
    integer :: idx = 0, num = 0, dim = 0, new_num_up = 0, new_num_down = 0
    type(Sub_Basis_Map), allocatable :: sub(:)
  end type Basis_Map

  type Total_Basis
    integer :: len = 0, num = 0, dim = 0
    type(Basis_Map), allocatable :: sub(:)
  end type Total_Basis

  type Sub_Super_Basis
    integer :: sys_num_up = 0, sys_num_down = 0, env_num_up = 0, env_num_down = 0, sys_dim = 0, env_dim = 0
  end type Sub_Super_Basis

  type Super_Basis
    integer :: sys_len = 0, env_len = 0, num = 0, dim = 0, idx = 0
    type(Sub_Super_Basis), allocatable :: sub(:)
  end type Super_Basis

  type Sub_Wave
    integer :: sys_num_up = 0, sys_num_down = 0, env_num_up = 0, env_num_down = 0, sys_dim = 0, env_dim = 0
    complex(c_double_complex), allocatable :: vec(:, :)
  end type Sub_Wave

  type Wavefunction
    integer :: sys_len = 0, env_len = 0, num = 0, dim = 0
    type(Sub_Wave), allocatable :: sub(:)
  end type Wavefunction

  ! globals the reference keeps in pubdata (common.f90:66-71)
  integer :: num_site = 0, tot_num_up = 0, tot_num_down = 0
  integer :: kept_min = 110, kept_max = 300
  real(8) :: lanzero = 1.0d-6, Entropys = 0.0d0, Energys(8) = 0.0d0, Energy0 = 0.0d0
  ! H.psi precision of lanczos_memo_new (TSC_PREC_*) and its FP64 refinement steps
  integer(c_int) :: lanczos_prec = 0, lanczos_refine = 0
  ! lattice bonds
  type(tsc_bond), allocatable, target, private :: bonds(:)
  type(c_ptr), save, private :: ctx = c_null_ptr

  ! packed block model: C descriptors and the arrays they point at
  type, private :: OpPack
    type(tsc_op_block), allocatable :: b(:)
  end type OpPack

  type, private :: ModelPack
    type(tsc_sector), allocatable :: sec(:)
    integer(c_int32_t), allocatable :: sites(:)
    type(OpPack), allocatable :: ops(:)
    type(tsc_op_desc), allocatable :: desc(:)
    type(tsc_site_ops), allocatable :: keep(:)
    type(tsc_model_desc) :: d
  end type ModelPack

  private :: pack_model, arena_add, dev, sector_of, flatten_wave, unflatten_wave, superblock_plan, find_trun, &
             env_signs, env_sector

contains

  ! ---- set-up --------------------------------------------------------------

  subroutine dmrg_setup_bonds(nsites, n, bi, bj, bt, bjex)
    integer, intent(in) :: nsites, n, bi(n), bj(n)
    real(8), intent(in) :: bt(n), bjex(n)
    integer :: k
    num_site = nsites
    if (allocated(bonds)) deallocate(bonds)
    allocate(bonds(max(n, 1)))
    do k = 1, n
      bonds(k) = tsc_bond(int(bi(k), c_int32_t), int(bj(k), c_int32_t), real(bt(k), c_double), real(bjex(k), c_double))
    end do
    if (n == 0) bonds(1) = tsc_bond(0, 0, 0.0d0, 0.0d0)
  end subroutine dmrg_setup_bonds

  integer function nbonds()
    nbonds = 0
    if (allocated(bonds)) then
      nbonds = size(bonds)
      if (nbonds == 1 .and. bonds(1)%i == 0) nbonds = 0
    end if
  end function nbonds

  type(c_ptr) function dev()
    if (.not. c_associated(ctx)) call tsc_check(tsc_ctx_create(0_c_int, ctx), 'tsc_ctx_create')
    dev = ctx
  end function dev

  ! Get_hamiltonian_site (dmrg_sub.f90:913-942) with the site operators of
  ! Get_site_operator_su2 (dmrg_sub.f90:733-908): c+ = -sqrt(2) hole ->
  ! electron, S = sqrt(3/2) and n = sqrt(2) on the electron
  subroutine Get_hamiltonian_site(site, blk)
    integer, intent(in) :: site
    type(Total_Model), intent(inout) :: blk
    call deallocate_model(blk)
    blk%len = 1
    allocate(blk%sites(1), blk%keep(1), blk%sub(1))
    blk%sites(1) = site
    blk%keep(1) = site
    call new_block(blk%ham, 0, 2)
    call set_sub(blk%ham%sub(1), 1, 1, 0, reshape([(0.0d0, 0.0d0)], [1, 1]))
    call set_sub(blk%ham%sub(2), 0, 0, 0, reshape([(0.0d0, 0.0d0)], [1, 1]))
    call new_block(blk%sub(1)%elec_up, 1, 1)
    blk%sub(1)%elec_up%down_dif = 1
    call set_sub(blk%sub(1)%elec_up%sub(1), 0, 0, 1, reshape([cmplx(-sqrt(2.0d0), 0.0d0, c_double_complex)], [1, 1]))
    call new_block(blk%sub(1)%spin_sd, 0, 1)
    blk%sub(1)%spin_sd%down_dif = 2
    call set_sub(blk%sub(1)%spin_sd%sub(1), 1, 1, 0, reshape([cmplx(sqrt(1.5d0), 0.0d0, c_double_complex)], [1, 1]))
    call new_block(blk%sub(1)%num_sn, 0, 1)
    call set_sub(blk%sub(1)%num_sn%sub(1), 1, 1, 0, reshape([cmplx(sqrt(2.0d0), 0.0d0, c_double_complex)], [1, 1]))
  end subroutine Get_hamiltonian_site

  subroutine new_block(b, up_dif, num)
    type(Total_Block), intent(inout) :: b
    integer, intent(in) :: up_dif, num
    if (allocated(b%sub)) deallocate(b%sub)
    b%up_dif = up_dif
    b%num = num
    allocate(b%sub(num))
  end subroutine new_block

  subroutine set_sub(s, num_up, num_down, down_dif, m)
    type(Sub_Block), intent(inout) :: s
    integer, intent(in) :: num_up, num_down, down_dif
    complex(c_double_complex), intent(in) :: m(:, :)
    s%num_up = num_up
    s%num_down = num_down
    s%down_dif = down_dif
    s%row_dim = size(m, 1)
    s%sdim = size(m, 2)
    if (allocated(s%mat)) deallocate(s%mat)
    allocate(s%mat(s%row_dim, s%sdim))
    s%mat = m
  end subroutine set_sub

  subroutine deallocate_block(b)
    type(Total_Block), intent(inout) :: b
    if (allocated(b%sub)) deallocate(b%sub)
    b%num = 0
  end subroutine deallocate_block

  subroutine deallocate_model(m)
    type(Total_Model), intent(inout) :: m
    call deallocate_block(m%ham)
    if (allocated(m%sites)) deallocate(m%sites)
    if (allocated(m%keep)) deallocate(m%keep)
    if (allocated(m%sub)) deallocate(m%sub)
    m%len = 0
  end subroutine deallocate_model

  ! ---- bases (block.f90:2087-2245) -----------------------------------------

  ! enlarged basis of block (x) site: (N, 2S) pairs sorted by 2S then N
  ! descending (QuickSort_Integer's two exchange passes), unique, and the
  ! (block sector, site state) sub-blocks of each in block x site order
  subroutine Get_basis(ham, basis)
    type(Total_Block), intent(in) :: ham
    type(Total_Basis), intent(inout) :: basis
    integer, parameter :: sn(2) = [1, 0], sj(2) = [1, 0]
    integer, allocatable :: pn(:), pj(:)
    integer :: a, s, j, n, i, x, m, t, cnt, nu
    n = 0
    do a = 1, ham%num
      do s = 1, 2
        n = n + (ham%sub(a)%num_down + sj(s) - abs(ham%sub(a)%num_down - sj(s))) / 2 + 1
      end do
    end do
    allocate(pn(n), pj(n))
    n = 0
    do a = 1, ham%num
      do s = 1, 2
        do j = abs(ham%sub(a)%num_down - sj(s)), ham%sub(a)%num_down + sj(s), 2
          n = n + 1
          pn(n) = ham%sub(a)%num_up + sn(s)
          pj(n) = j
        end do
      end do
    end do
    do i = 1, n
      do x = i + 1, n
        if (pj(i) < pj(x)) then
          t = pj(i); pj(i) = pj(x); pj(x) = t
          t = pn(i); pn(i) = pn(x); pn(x) = t
        end if
      end do
    end do
    do i = 1, n
      do x = i + 1, n
        if (pj(i) == pj(x) .and. pn(i) < pn(x)) then
          t = pn(i); pn(i) = pn(x); pn(x) = t
        end if
      end do
    end do
    nu = 0
    do i = 1, n
      if (i > 1) then
        if (pn(i) == pn(i - 1) .and. pj(i) == pj(i - 1)) cycle
      end if
      nu = nu + 1
      pn(nu) = pn(i)
      pj(nu) = pj(i)
    end do
    if (allocated(basis%sub)) deallocate(basis%sub)
    allocate(basis%sub(nu))
    basis%num = nu
    basis%dim = 0
    do x = 1, nu
      basis%sub(x)%idx = x
      basis%sub(x)%new_num_up = pn(x)
      basis%sub(x)%new_num_down = pj(x)
      cnt = 0
      do a = 1, ham%num
        do s = 1, 2
          if (ham%sub(a)%num_up + sn(s) == pn(x) .and. abs(ham%sub(a)%num_down - sj(s)) <= pj(x) .and. &
              pj(x) <= ham%sub(a)%num_down + sj(s)) cnt = cnt + 1
        end do
      end do
      allocate(basis%sub(x)%sub(cnt))
      basis%sub(x)%num = cnt
      m = 0
      cnt = 0
      do a = 1, ham%num
        do s = 1, 2
          if (ham%sub(a)%num_up + sn(s) == pn(x) .and. abs(ham%sub(a)%num_down - sj(s)) <= pj(x) .and. &
              pj(x) <= ham%sub(a)%num_down + sj(s)) then
            cnt = cnt + 1
            basis%sub(x)%sub(cnt) = Sub_Basis_Map(m, ham%sub(a)%sdim, ham%sub(a)%num_up, ham%sub(a)%num_down, &
                                                  sn(s), sj(s))
            m = m + ham%sub(a)%sdim
          end if
        end do
      end do
      basis%sub(x)%dim = m
      basis%dim = basis%dim + m
    end do
  end subroutine Get_basis

  ! sys sectors outer, env inner, N adding up and 2S coupling to s2tot
  subroutine Get_super_basis(sys_bs, env_bs, super, ntot, s2tot)
    type(Total_Basis), intent(in) :: sys_bs, env_bs
    type(Super_Basis), intent(inout) :: super
    integer, intent(in) :: ntot, s2tot
    integer :: x, y, n, pass
    do pass = 1, 2
      n = 0
      super%dim = 0
      do x = 1, sys_bs%num
        do y = 1, env_bs%num
          if (env_bs%sub(y)%new_num_up /= ntot - sys_bs%sub(x)%new_num_up) cycle
          if (abs(sys_bs%sub(x)%new_num_down - env_bs%sub(y)%new_num_down) > s2tot) cycle
          if (s2tot > sys_bs%sub(x)%new_num_down + env_bs%sub(y)%new_num_down) cycle
          n = n + 1
          super%dim = super%dim + sys_bs%sub(x)%dim * env_bs%sub(y)%dim
          if (pass == 2) super%sub(n) = Sub_Super_Basis(sys_bs%sub(x)%new_num_up, sys_bs%sub(x)%new_num_down, &
                                                         env_bs%sub(y)%new_num_up, env_bs%sub(y)%new_num_down, &
                                                         sys_bs%sub(x)%dim, env_bs%sub(y)%dim)
        end do
      end do
      if (pass == 1) then
        if (allocated(super%sub)) deallocate(super%sub)
        allocate(super%sub(n))
        super%num = n
      end if
    end do
  end subroutine Get_super_basis

  subroutine allocate_wave(super, wave)
    type(Super_Basis), intent(in) :: super
    type(Wavefunction), intent(inout) :: wave
    integer :: i
    if (allocated(wave%sub)) deallocate(wave%sub)
    allocate(wave%sub(super%num))
    wave%sys_len = super%sys_len
    wave%env_len = super%env_len
    wave%num = super%num
    wave%dim = super%dim
    do i = 1, super%num
      wave%sub(i)%sys_num_up = super%sub(i)%sys_num_up
      wave%sub(i)%sys_num_down = super%sub(i)%sys_num_down
      wave%sub(i)%env_num_up = super%sub(i)%env_num_up
      wave%sub(i)%env_num_down = super%sub(i)%env_num_down
      wave%sub(i)%sys_dim = super%sub(i)%sys_dim
      wave%sub(i)%env_dim = super%sub(i)%env_dim
      allocate(wave%sub(i)%vec(super%sub(i)%sys_dim, super%sub(i)%env_dim))
      wave%sub(i)%vec = (0.0d0, 0.0d0)
    end do
  end subroutine allocate_wave

  ! wave_initialize (block.f90:1154-1180): uniform [0, 1) real and imaginary
  ! parts, normalised; seeded (the reference's random_number is not)
  subroutine wave_initialize(wave, seed)
    type(Wavefunction), intent(inout) :: wave
    integer, intent(in) :: seed
    integer :: i, j, k
    integer(8) :: st
    real(8) :: re, im, nrm
    st = int(seed, 8) * 2654435761_8 + 12345_8
    nrm = 0.0d0
    do i = 1, wave%num
      do k = 1, wave%sub(i)%env_dim
        do j = 1, wave%sub(i)%sys_dim
          re = lcg(st)
          im = lcg(st)
          wave%sub(i)%vec(j, k) = cmplx(re, im, c_double_complex)
          nrm = nrm + re * re + im * im
        end do
      end do
    end do
    do i = 1, wave%num
      wave%sub(i)%vec = wave%sub(i)%vec / sqrt(nrm)
    end do
  contains
    real(8) function lcg(s)
      integer(8), intent(inout) :: s
      s = iand(s * 6364136223846793005_8 + 1442695040888963407_8, huge(s))
      lcg = real(ishft(s, -11), 8) / 4503599627370496.0d0
    end function lcg
  end subroutine wave_initialize

  ! ---- packing ---------------------------------------------------------------

  subroutine arena_add(ar, n, m, conj)
    complex(c_double_complex), allocatable, intent(inout) :: ar(:)
    integer(8), intent(inout) :: n
    complex(c_double_complex), intent(in) :: m(:, :)
    logical, intent(in) :: conj
    complex(c_double_complex), allocatable :: tmp(:)
    integer(8) :: need
    need = n + size(m, kind=8)
    if (.not. allocated(ar)) allocate(ar(max(need, 1024_8)))
    if (need > size(ar, kind=8)) then
      allocate(tmp(max(need, 2 * size(ar, kind=8))))
      tmp(1:n) = ar(1:n)
      call move_alloc(tmp, ar)
    end if
    if (conj) then
      ar(n + 1:need) = reshape(conjg(m), [size(m, kind=8)])
    else
      ar(n + 1:need) = reshape(m, [size(m, kind=8)])
    end if
    n = need
  end subroutine arena_add

  integer function sector_of(ham, num_up, num_down)
    type(Total_Block), intent(in) :: ham
    integer, intent(in) :: num_up, num_down
    integer :: a
    sector_of = -1
    do a = 1, ham%num
      if (ham%sub(a)%num_up == num_up .and. ham%sub(a)%num_down == num_down) then
        sector_of = a - 1
        return
      end if
    end do
  end function sector_of

  ! C descriptors of a block model; its matrices appended to the arena
  ! (conjugated for an env block, whose operators the types hold as the
  ! reference stores them).
  subroutine pack_model(m, side, ar, n, pk, conj)
    type(Total_Model), intent(in) :: m
    integer, intent(in) :: side
    logical, intent(in) :: conj
    complex(c_double_complex), allocatable, intent(inout) :: ar(:)
    integer(8), intent(inout) :: n
    type(ModelPack), target, intent(inout) :: pk
    integer :: a, k, w, nk
    nk = 0
    if (allocated(m%keep)) nk = size(m%keep)
    allocate(pk%sec(max(m%ham%num, 1)), pk%sites(max(m%len, 1)), pk%ops(1 + 3 * nk), pk%desc(1 + 3 * nk), &
             pk%keep(max(nk, 1)))
    do a = 1, m%ham%num
      pk%sec(a) = tsc_sector(m%ham%sub(a)%num_up, m%ham%sub(a)%num_down, m%ham%sub(a)%sdim)
    end do
    pk%sites(1:m%len) = m%sites(1:m%len)
    call op(m%ham, 0, 0, 1)
    do k = 1, nk
      do w = 1, 3
        select case (w)
        case (1)
          call op(m%sub(k)%elec_up, 1, 1, 1 + 3 * (k - 1) + w)
        case (2)
          call op(m%sub(k)%spin_sd, 0, 2, 1 + 3 * (k - 1) + w)
        case (3)
          call op(m%sub(k)%num_sn, 0, 0, 1 + 3 * (k - 1) + w)
        end select
      end do
      pk%keep(k)%site = m%keep(k)
      pk%keep(k)%op_c = c_loc(pk%desc(1 + 3 * (k - 1) + 1))
      pk%keep(k)%op_s = c_loc(pk%desc(1 + 3 * (k - 1) + 2))
      pk%keep(k)%op_n = c_loc(pk%desc(1 + 3 * (k - 1) + 3))
    end do
    pk%d = tsc_model_desc(int(m%ham%num, c_int32_t), c_loc(pk%sec), int(m%len, c_int32_t), c_loc(pk%sites), &
                          c_loc(pk%desc(1)), int(nk, c_int32_t), c_loc(pk%keep))
  contains
    subroutine op(b, up_dif, k2, slot)
      type(Total_Block), intent(in) :: b
      integer, intent(in) :: up_dif, k2, slot
      integer :: q, bra, ket
      allocate(pk%ops(slot)%b(max(b%num, 1)))
      do q = 1, b%num
        ket = sector_of(m%ham, b%sub(q)%num_up, b%sub(q)%num_down)
        bra = sector_of(m%ham, b%sub(q)%num_up + b%up_dif, b%sub(q)%num_down + b%sub(q)%down_dif)
        if (ket < 0 .or. bra < 0) error stop 'operator sub-block outside the block sectors'
        pk%ops(slot)%b(q) = tsc_op_block(int(bra, c_int32_t), int(ket, c_int32_t), int(n, c_int64_t))
        call arena_add(ar, n, b%sub(q)%mat, conj)
      end do
      pk%desc(slot) = tsc_op_desc(int(side, c_int32_t), int(up_dif, c_int32_t), int(k2, c_int32_t), &
                                  int(b%num, c_int32_t), c_loc(pk%ops(slot)%b))
    end subroutine op
  end subroutine pack_model

  ! (-1)^(N_block N_site) over the rows of sector x of an enlarged env
  ! basis: -1 where the product state has an odd block and an odd site
  subroutine env_signs(basis, x, sg)
    type(Total_Basis), intent(in) :: basis
    integer, intent(in) :: x
    real(8), allocatable, intent(out) :: sg(:)
    integer :: j, p
    allocate(sg(max(basis%sub(x)%dim, 1)))
    sg = 1.0d0
    do j = 1, basis%sub(x)%num
      if (mod(basis%sub(x)%sub(j)%bl_num_up, 2) /= 0 .and. mod(basis%sub(x)%sub(j)%st_num_up, 2) /= 0) then
        p = basis%sub(x)%sub(j)%spos
        sg(p + 1:p + basis%sub(x)%sub(j)%sdim) = -1.0d0
      end if
    end do
  end subroutine env_signs

  integer function env_sector(basis, num_up, num_down)
    type(Total_Basis), intent(in) :: basis
    integer, intent(in) :: num_up, num_down
    integer :: x
    env_sector = 0
    do x = 1, basis%num
      if (basis%sub(x)%new_num_up == num_up .and. basis%sub(x)%new_num_down == num_down) env_sector = x
    end do
    if (env_sector == 0) error stop 'wave function sector outside the env basis'
  end function env_sector

  ! the library's layout of a wave function: sectors end to end, env
  ! columns times env_signs (the same map both ways)
  subroutine flatten_wave(w, env_bs, flat)
    type(Wavefunction), intent(in) :: w
    type(Total_Basis), intent(in) :: env_bs
    complex(c_double_complex), intent(out) :: flat(:)
    real(8), allocatable :: sg(:)
    integer(8) :: o
    integer :: i, c
    o = 0
    do i = 1, w%num
      call env_signs(env_bs, env_sector(env_bs, w%sub(i)%env_num_up, w%sub(i)%env_num_down), sg)
      do c = 1, w%sub(i)%env_dim
        flat(o + 1:o + w%sub(i)%sys_dim) = w%sub(i)%vec(:, c) * sg(c)
        o = o + w%sub(i)%sys_dim
      end do
    end do
  end subroutine flatten_wave

  subroutine unflatten_wave(flat, env_bs, w)
    complex(c_double_complex), intent(in) :: flat(:)
    type(Total_Basis), intent(in) :: env_bs
    type(Wavefunction), intent(inout) :: w
    real(8), allocatable :: sg(:)
    integer(8) :: o
    integer :: i, c
    o = 0
    do i = 1, w%num
      call env_signs(env_bs, env_sector(env_bs, w%sub(i)%env_num_up, w%sub(i)%env_num_down), sg)
      do c = 1, w%sub(i)%env_dim
        w%sub(i)%vec(:, c) = flat(o + 1:o + w%sub(i)%sys_dim) * sg(c)
        o = o + w%sub(i)%sys_dim
      end do
    end do
  end subroutine unflatten_wave

  ! the device plan of a superblock [sys, s1 | env, s2]
  subroutine superblock_plan(sys, env, ntot, plan, dim)
    type(Total_Model), intent(in) :: sys, env
    integer, intent(in) :: ntot
    type(c_ptr), intent(out) :: plan
    integer(8), intent(out) :: dim
    type(ModelPack), target :: ps, pe
    complex(c_double_complex), allocatable, target :: ar(:)
    integer(8) :: n
    type(c_ptr) :: tasks
    integer(c_int64_t), target :: cnt(9)
    n = 0
    call pack_model(sys, 0, ar, n, ps, .false.)
    call pack_model(env, 1, ar, n, pe, .true.)
    call tsc_check(tsc_model_tasks(c_loc(ps%d), c_loc(pe%d), int(sys%len + 1, c_int32_t), &
                                   int(num_site - env%len, c_int32_t), int(nbonds(), c_int32_t), c_loc(bonds), &
                                   int(ntot, c_int32_t), int(tot_num_down, c_int32_t), c_loc(ar), &
                                   int(n, c_int64_t), tasks), 'tsc_model_tasks')
    call tsc_check(tsc_tasks_info(tasks, cnt), 'tsc_tasks_info')
    dim = cnt(6)
    call tsc_check(tsc_tasks_plan(dev(), tasks, c_null_ptr, 0_c_int64_t, plan), 'tsc_tasks_plan')
    call tsc_check(tsc_tasks_destroy(tasks), 'tsc_tasks_destroy')
  end subroutine superblock_plan

  ! ---- the drop-in subroutines ---------------------------------------------

  ! hmerge (get_sys.f90:7398-7626) forms the merged partner operators of
  ! every boundary class before model_wave; here tsc_model_tasks forms them
  ! inside each plan, so the blocks are left as they are.
  subroutine hmerge(sys, env, sys_bs, env_bs)
    type(Total_Model), intent(inout) :: sys, env
    type(Total_Basis), intent(in) :: sys_bs, env_bs
    if (sys%len < 0 .or. env%len < 0 .or. sys_bs%num < 0 .or. env_bs%num < 0) error stop 'hmerge: bad blocks'
  end subroutine hmerge

  ! outvec = H invec on the superblock [sys, s1 | env, s2] (outvec allocated
  ! by the caller with allocate_wave, overwritten)
  subroutine model_wave(sys, env, sys_bs, env_bs, invec, outvec)
    type(Total_Model), intent(in) :: sys, env
    type(Total_Basis), intent(in) :: sys_bs, env_bs
    type(Wavefunction), intent(in) :: invec
    type(Wavefunction), intent(inout) :: outvec
    type(c_ptr) :: plan
    integer(8) :: dim
    complex(c_double_complex), allocatable, target :: xin(:), xout(:)
    if (sys_bs%num < 1 .or. env_bs%num < 1) error stop 'model_wave: empty basis'
    call superblock_plan(sys, env, invec%sub(1)%sys_num_up + invec%sub(1)%env_num_up, plan, dim)
    if (dim /= invec%dim) error stop 'model_wave: superblock dimension differs from the wavefunction'
    allocate(xin(dim), xout(dim))
    call flatten_wave(invec, env_bs, xin)
    call tsc_check(tsc_hpsi_host(plan, c_loc(xin), c_loc(xout), TSC_PREC_F64), 'tsc_hpsi_host')
    call unflatten_wave(xout, env_bs, outvec)
    call tsc_check(tsc_plan_destroy(plan), 'tsc_plan_destroy')
  end subroutine model_wave

  ! lowest eigenpair (levels = 1): eigvec(1) holds the start vector on entry
  ! and the Ritz vector on exit; Energys(1) and Energy0 are set
  subroutine lanczos_memo_new(sys, env, sys_bs, env_bs, super, eigvec, levels, num_idx)
    type(Total_Model), intent(in) :: sys, env
    type(Total_Basis), intent(in) :: sys_bs, env_bs
    type(Super_Basis), intent(in) :: super
    integer, intent(in) :: levels, num_idx
    type(Wavefunction), intent(inout) :: eigvec(levels)
    type(c_ptr) :: plan, psi
    integer(8) :: dim
    real(c_double) :: e
    integer(c_int32_t) :: info(2)
    complex(c_double_complex), allocatable, target :: x(:)
    if (levels /= 1) error stop 'lanczos_memo_new: levels = 1 only'
    if (sys_bs%num < 1 .or. env_bs%num < 1 .or. num_idx < 0) error stop 'lanczos_memo_new: bad arguments'
    call superblock_plan(sys, env, super%sub(1)%sys_num_up + super%sub(1)%env_num_up, plan, dim)
    if (dim /= super%dim) error stop 'lanczos_memo_new: superblock dimension differs from the super basis'
    allocate(x(dim))
    call flatten_wave(eigvec(1), env_bs, x)
    call tsc_check(tsc_vec_create(dev(), int(dim, c_int64_t), psi), 'tsc_vec_create')
    call tsc_check(tsc_vec_upload(psi, c_loc(x)), 'tsc_vec_upload')
    call tsc_check(tsc_lanczos(plan, psi, real(lanzero, c_double), 100_c_int, 4_c_int, lanczos_prec, &
                               lanczos_refine, e, info), 'tsc_lanczos')
    call tsc_check(tsc_vec_download(psi, c_loc(x)), 'tsc_vec_download')
    call unflatten_wave(x, env_bs, eigvec(1))
    Energys(1) = e
    Energy0 = e
    call tsc_check(tsc_vec_destroy(psi), 'tsc_vec_destroy')
    call tsc_check(tsc_plan_destroy(plan), 'tsc_plan_destroy')
  end subroutine lanczos_memo_new

  subroutine density_matrix(den, wave, side)
    type(Total_Block), intent(inout) :: den
    type(Wavefunction), intent(in) :: wave
    integer, intent(in) :: side
    integer(c_int32_t), allocatable, target :: sd(:), ed(:)
    complex(c_double_complex), allocatable, target :: x(:), rho(:)
    type(c_ptr) :: psi
    integer :: i, d
    integer(8) :: o, nr
    allocate(sd(wave%num), ed(wave%num), x(max(wave%dim, 1)))
    nr = 0
    do i = 1, wave%num
      sd(i) = wave%sub(i)%sys_dim
      ed(i) = wave%sub(i)%env_dim
      d = merge(wave%sub(i)%sys_dim, wave%sub(i)%env_dim, side == 0)
      nr = nr + int(d, 8) * d
    end do
    allocate(rho(max(nr, 1_8)))
    ! rho of the wave function as the types hold it (the reference's
    ! fermion order): the reference's rho_sys = psi psi^H, rho_env = psi^H psi
    o = 0
    do i = 1, wave%num
      x(o + 1:o + size(wave%sub(i)%vec, kind=8)) = reshape(wave%sub(i)%vec, [size(wave%sub(i)%vec, kind=8)])
      o = o + size(wave%sub(i)%vec, kind=8)
    end do
    call tsc_check(tsc_vec_create(dev(), int(wave%dim, c_int64_t), psi), 'tsc_vec_create')
    call tsc_check(tsc_vec_upload(psi, c_loc(x)), 'tsc_vec_upload')
    call tsc_check(tsc_density_matrix(dev(), psi, int(wave%num, c_int32_t), c_loc(sd), c_loc(ed), &
                                      int(side, c_int), c_loc(rho)), 'tsc_density_matrix')
    call tsc_check(tsc_vec_destroy(psi), 'tsc_vec_destroy')
    call new_block(den, 0, wave%num)
    o = 0
    do i = 1, wave%num
      d = merge(wave%sub(i)%sys_dim, wave%sub(i)%env_dim, side == 0)
      if (side == 0) then
        call set_sub(den%sub(i), wave%sub(i)%sys_num_up, wave%sub(i)%sys_num_down, 0, &
                     reshape(rho(o + 1:o + int(d, 8) * d), [d, d]))
      else
        call set_sub(den%sub(i), wave%sub(i)%env_num_up, wave%sub(i)%env_num_down, 0, &
                     reshape(rho(o + 1:o + int(d, 8) * d), [d, d]))
      end if
      o = o + int(d, 8) * d
    end do
  end subroutine density_matrix

  ! rho_s = psi_s psi_s^H per psi sector (den_mat.f90:4-54)
  subroutine Get_density_matrix_sys(sys_den, wave)
    type(Total_Block), intent(inout) :: sys_den
    type(Wavefunction), intent(in) :: wave
    call density_matrix(sys_den, wave, 0)
  end subroutine Get_density_matrix_sys

  ! rho_s = psi_s^H psi_s per psi sector (den_mat.f90:55-105)
  subroutine Get_density_matrix_env(env_den, wave)
    type(Total_Block), intent(inout) :: env_den
    type(Wavefunction), intent(in) :: wave
    call density_matrix(env_den, wave, 1)
  end subroutine Get_density_matrix_env

  ! eigensolve of every rho sector, the kept-state rule and the truncation
  ! operator: trun_op%sub = the top kept eigenvectors of each sector with
  ! kept > 0 (row_dim = sector dim, sdim = kept); Entropys is set
  subroutine Get_truncation_operator(den_mat, trun_op, trun_err, num_state)
    type(Total_Block), intent(in) :: den_mat
    type(Total_Block), intent(inout) :: trun_op
    real(8), intent(in) :: trun_err
    integer, intent(out) :: num_state
    integer(c_int32_t), allocatable, target :: dim(:), twos(:), kept(:)
    complex(c_double_complex), allocatable, target :: rho(:), u(:)
    real(c_double), allocatable, target :: eig(:)
    real(c_double), target :: scal(4)
    integer :: i, nk, q
    integer(8) :: o, ucap, nr, ne
    allocate(dim(max(den_mat%num, 1)), twos(max(den_mat%num, 1)), kept(max(den_mat%num, 1)))
    nr = 0
    ne = 0
    do i = 1, den_mat%num
      dim(i) = den_mat%sub(i)%sdim
      twos(i) = den_mat%sub(i)%num_down
      nr = nr + int(dim(i), 8) * dim(i)
      ne = ne + dim(i)
    end do
    ucap = max(nr, 1_8)
    allocate(rho(max(nr, 1_8)), u(ucap), eig(max(ne, 1_8)))
    o = 0
    do i = 1, den_mat%num
      rho(o + 1:o + int(dim(i), 8) * dim(i)) = reshape(den_mat%sub(i)%mat, [int(dim(i), 8) * dim(i)])
      o = o + int(dim(i), 8) * dim(i)
    end do
    call tsc_check(tsc_truncation_operator(dev(), int(den_mat%num, c_int32_t), c_loc(dim), c_loc(twos), c_loc(rho), &
                                           int(kept_min, c_int32_t), int(kept_max, c_int32_t), &
                                           real(trun_err, c_double), c_loc(eig), c_loc(kept), c_loc(u), &
                                           int(ucap, c_int64_t), c_loc(scal)), 'tsc_truncation_operator')
    num_state = nint(scal(1))
    Entropys = scal(3)
    nk = count(kept(1:den_mat%num) > 0)
    call new_block(trun_op, 0, nk)
    o = 0
    q = 0
    do i = 1, den_mat%num
      if (kept(i) == 0) cycle
      q = q + 1
      call set_sub(trun_op%sub(q), den_mat%sub(i)%num_up, den_mat%sub(i)%num_down, 0, &
                   reshape(u(o + 1:o + int(dim(i), 8) * kept(i)), [int(dim(i)), int(kept(i))]))
      o = o + int(dim(i), 8) * kept(i)
    end do
  end subroutine Get_truncation_operator

  integer function find_trun(trun, num_up, num_down)
    type(Total_Block), intent(in) :: trun
    integer, intent(in) :: num_up, num_down
    integer :: k
    find_trun = 0
    do k = 1, trun%num
      if (trun%sub(k)%num_up == num_up .and. trun%sub(k)%num_down == num_down) then
        find_trun = k
        return
      end if
    end do
  end function find_trun

  ! eff%sub = U_bra^H ori%sub U_ket for every sub-block whose bra and ket
  ! sectors the truncation keeps (block.f90:3465-3683)
  subroutine update_trun(ori, eff, trun)
    type(Total_Block), intent(in) :: ori, trun
    type(Total_Block), intent(inout) :: eff
    integer(c_int32_t), allocatable, target :: row(:), col(:), kl(:), kr(:)
    integer(c_int64_t), allocatable, target :: oo(:), lo(:), ro(:), eo(:)
    complex(c_double_complex), allocatable, target :: ob(:), ub(:), eb(:)
    integer, allocatable :: pick(:), tb(:), tk(:)
    integer :: k, n, q
    integer(8) :: no, nu, ne
    allocate(pick(max(ori%num, 1)), tb(max(ori%num, 1)), tk(max(ori%num, 1)))
    n = 0
    do k = 1, ori%num
      q = find_trun(trun, ori%sub(k)%num_up + ori%up_dif, ori%sub(k)%num_down + ori%sub(k)%down_dif)
      if (q == 0) cycle
      if (find_trun(trun, ori%sub(k)%num_up, ori%sub(k)%num_down) == 0) cycle
      n = n + 1
      pick(n) = k
      tb(n) = q
      tk(n) = find_trun(trun, ori%sub(k)%num_up, ori%sub(k)%num_down)
    end do
    allocate(row(max(n, 1)), col(max(n, 1)), kl(max(n, 1)), kr(max(n, 1)), oo(max(n, 1)), lo(max(n, 1)), &
             ro(max(n, 1)), eo(max(n, 1)))
    no = 0
    nu = 0
    ne = 0
    do q = 1, n
      k = pick(q)
      row(q) = ori%sub(k)%row_dim
      col(q) = ori%sub(k)%sdim
      kl(q) = trun%sub(tb(q))%sdim
      kr(q) = trun%sub(tk(q))%sdim
      if (trun%sub(tb(q))%row_dim /= row(q) .or. trun%sub(tk(q))%row_dim /= col(q)) &
        error stop 'update_trun: truncation and operator dimensions differ'
      oo(q) = no
      no = no + int(row(q), 8) * col(q)
      lo(q) = nu
      nu = nu + int(row(q), 8) * kl(q)
      ro(q) = nu
      nu = nu + int(col(q), 8) * kr(q)
      eo(q) = ne
      ne = ne + int(kl(q), 8) * kr(q)
    end do
    allocate(ob(max(no, 1_8)), ub(max(nu, 1_8)), eb(max(ne, 1_8)))
    do q = 1, n
      k = pick(q)
      ob(oo(q) + 1:oo(q) + int(row(q), 8) * col(q)) = reshape(ori%sub(k)%mat, [int(row(q), 8) * col(q)])
      ub(lo(q) + 1:lo(q) + int(row(q), 8) * kl(q)) = reshape(trun%sub(tb(q))%mat, [int(row(q), 8) * kl(q)])
      ub(ro(q) + 1:ro(q) + int(col(q), 8) * kr(q)) = reshape(trun%sub(tk(q))%mat, [int(col(q), 8) * kr(q)])
    end do
    if (n > 0) call tsc_check(tsc_update_trun(dev(), int(n, c_int32_t), c_loc(row), c_loc(col), c_loc(kl), &
                                              c_loc(kr), c_loc(oo), c_loc(lo), c_loc(ro), c_loc(eo), c_loc(ob), &
                                              int(no, c_int64_t), c_loc(ub), int(nu, c_int64_t), c_loc(eb), &
                                              int(ne, c_int64_t), TSC_PREC_F64), 'tsc_update_trun')
    call new_block(eff, ori%up_dif, n)
    eff%len = ori%len + 1
    do q = 1, n
      k = pick(q)
      call set_sub(eff%sub(q), ori%sub(k)%num_up, ori%sub(k)%num_down, ori%sub(k)%down_dif, &
                   reshape(eb(eo(q) + 1:eo(q) + int(kl(q), 8) * kr(q)), [int(kl(q)), int(kr(q))]))
    end do
  end subroutine update_trun

  subroutine update_trun_dia(ori, eff, trun)
    type(Total_Block), intent(in) :: ori, trun
    type(Total_Block), intent(inout) :: eff
    call update_trun(ori, eff, trun)
  end subroutine update_trun_dia

  subroutine update_trun_ndia(ori, eff, trun)
    type(Total_Block), intent(in) :: ori, trun
    type(Total_Block), intent(inout) :: eff
    call update_trun(ori, eff, trun)
  end subroutine update_trun_ndia

  ! pre (block) (x) new site, truncated by trun, into pre; next is left
  ! without operators (the enlarged block is never formed: embedding and
  ! U^H . U run as one plan)
  subroutine renormalize(pre, next, basis, trun, new_site, env_side)
    type(Total_Model), intent(inout) :: pre, next
    type(Total_Basis), intent(in) :: basis
    type(Total_Block), intent(in) :: trun
    integer, intent(in) :: new_site
    logical, intent(in) :: env_side
    type(ModelPack), target :: pk
    complex(c_double_complex), allocatable, target :: ar(:), uin(:), res(:)
    integer(c_int32_t), allocatable, target :: keep(:), order(:), kept(:), lay(:), nsec(:)
    integer(c_int64_t), allocatable, target :: offs(:)
    integer(c_int64_t), target :: c3(3), lens(3)
    integer, allocatable :: newsites(:), cnt(:)
    real(8), allocatable :: sg(:)
    integer(8) :: n, o
    integer :: k, x, q, nk, w, key, b2, k2, nw, ns, s, i
    type(c_ptr) :: tasks, plan
    type(Total_Model) :: nb
    logical :: out
    n = 0
    call pack_model(pre, 0, ar, n, pk, env_side)
    allocate(newsites(pre%len + 1))
    newsites(1:pre%len) = pre%sites(1:pre%len)
    newsites(pre%len + 1) = new_site
    ! carried sites: those with a bond leaving the new block
    allocate(keep(pre%len + 1))
    nk = 0
    do i = 1, pre%len + 1
      out = .false.
      do k = 1, nbonds()
        if (bonds(k)%i == newsites(i) .and. .not. any(newsites == bonds(k)%j)) out = .true.
        if (bonds(k)%j == newsites(i) .and. .not. any(newsites == bonds(k)%i)) out = .true.
      end do
      if (out) then
        nk = nk + 1
        keep(nk) = newsites(i)
      end if
    end do
    allocate(order(max(trun%num, 1)), kept(max(trun%num, 1)))
    o = 0
    do k = 1, trun%num
      x = 0
      do i = 1, basis%num
        if (basis%sub(i)%new_num_up == trun%sub(k)%num_up .and. basis%sub(i)%new_num_down == trun%sub(k)%num_down) x = i
      end do
      if (x == 0) error stop 'Get_Hamiltonian: truncation sector not in the enlarged basis'
      if (trun%sub(k)%row_dim /= basis%sub(x)%dim) error stop 'Get_Hamiltonian: truncation rows differ from the basis'
      order(k) = x - 1
      kept(k) = trun%sub(k)%sdim
      o = o + size(trun%sub(k)%mat, kind=8)
    end do
    allocate(uin(max(o, 1_8)))
    o = 0
    do k = 1, trun%num
      if (env_side) then
        ! the library's env U: the conjugated eigenvectors of the reference's
        ! rho_env, rows in its own fermion order (times env_signs)
        call env_signs(basis, order(k) + 1, sg)
        do i = 1, trun%sub(k)%sdim
          uin(o + 1:o + trun%sub(k)%row_dim) = conjg(trun%sub(k)%mat(:, i)) * sg(1:trun%sub(k)%row_dim)
          o = o + trun%sub(k)%row_dim
        end do
      else
        uin(o + 1:o + size(trun%sub(k)%mat, kind=8)) = reshape(trun%sub(k)%mat, [size(trun%sub(k)%mat, kind=8)])
        o = o + size(trun%sub(k)%mat, kind=8)
      end if
    end do
    call tsc_check(tsc_renorm_tasks(c_loc(pk%d), int(new_site, c_int32_t), int(nbonds(), c_int32_t), c_loc(bonds), &
                                    int(nk, c_int32_t), c_loc(keep), int(trun%num, c_int32_t), c_loc(order), &
                                    c_loc(kept), c_loc(ar), int(n, c_int64_t), tasks), 'tsc_renorm_tasks')
    call tsc_check(tsc_tasks_lengths(tasks, c_loc(lens)), 'tsc_tasks_lengths')
    if (lens(1) /= o) error stop 'Get_Hamiltonian: truncation length differs from the plan input'
    call tsc_check(tsc_tasks_layout(tasks, c_loc(c3), c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr), 'layout')
    nw = int(c3(1))
    ns = int(c3(2))
    allocate(lay(3 * max(nw, 1)), offs(max(nw, 1)), nsec(3 * max(ns, 1)), res(max(lens(2), 1_c_int64_t)))
    call tsc_check(tsc_tasks_layout(tasks, c_loc(c3), c_loc(lay), c_loc(offs), c_loc(nsec), c_null_ptr), 'layout')
    call tsc_check(tsc_tasks_plan(dev(), tasks, c_null_ptr, 0_c_int64_t, plan), 'tsc_tasks_plan')
    call tsc_check(tsc_hpsi_host(plan, c_loc(uin), c_loc(res), TSC_PREC_F64), 'tsc_hpsi_host')
    call tsc_check(tsc_plan_destroy(plan), 'tsc_plan_destroy')
    call tsc_check(tsc_tasks_destroy(tasks), 'tsc_tasks_destroy')
    ! the new block: Hamiltonian sub-blocks and the carried operators
    nb%len = pre%len + 1
    allocate(nb%sites(nb%len), nb%keep(nk), nb%sub(nk), cnt(0:3 * nk))
    nb%sites = newsites
    nb%keep = keep(1:nk)
    cnt = 0
    do w = 1, nw
      cnt(lay(3 * w - 2) + 1) = cnt(lay(3 * w - 2) + 1) + 1
    end do
    call new_block(nb%ham, 0, cnt(0))
    do k = 1, nk
      call new_block(nb%sub(k)%elec_up, 1, cnt(3 * (k - 1) + 1))
      call new_block(nb%sub(k)%spin_sd, 0, cnt(3 * (k - 1) + 2))
      call new_block(nb%sub(k)%num_sn, 0, cnt(3 * (k - 1) + 3))
    end do
    cnt = 0
    do w = 1, nw
      key = lay(3 * w - 2)
      b2 = lay(3 * w - 1) + 1
      k2 = lay(3 * w) + 1
      cnt(key + 1) = cnt(key + 1) + 1
      s = cnt(key + 1)
      ! an env block's operators are stored as the reference stores them
      if (env_side) res(offs(w) + 1:offs(w) + int(nsec(3 * b2), 8) * nsec(3 * k2)) = &
        conjg(res(offs(w) + 1:offs(w) + int(nsec(3 * b2), 8) * nsec(3 * k2)))
      associate (mat => reshape(res(offs(w) + 1:offs(w) + int(nsec(3 * b2), 8) * nsec(3 * k2)), &
                                [int(nsec(3 * b2)), int(nsec(3 * k2))]))
        if (key < 0) then
          call set_sub(nb%ham%sub(s), nsec(3 * k2 - 2), nsec(3 * k2 - 1), nsec(3 * b2 - 1) - nsec(3 * k2 - 1), mat)
        else
          k = key / 3 + 1
          select case (mod(key, 3))
          case (0)
            call set_sub(nb%sub(k)%elec_up%sub(s), nsec(3 * k2 - 2), nsec(3 * k2 - 1), &
                         nsec(3 * b2 - 1) - nsec(3 * k2 - 1), mat)
          case (1)
            call set_sub(nb%sub(k)%spin_sd%sub(s), nsec(3 * k2 - 2), nsec(3 * k2 - 1), &
                         nsec(3 * b2 - 1) - nsec(3 * k2 - 1), mat)
          case (2)
            call set_sub(nb%sub(k)%num_sn%sub(s), nsec(3 * k2 - 2), nsec(3 * k2 - 1), &
                         nsec(3 * b2 - 1) - nsec(3 * k2 - 1), mat)
          end select
        end if
      end associate
    end do
    call sort_ham_by_sector(nb%ham, nsec, ns)
    call deallocate_model(next)
    next%len = nb%len
    call deallocate_model(pre)
    pre = nb
  end subroutine renormalize

  ! the block's sectors are its Hamiltonian's diagonal sub-blocks in the new
  ! sector order
  subroutine sort_ham_by_sector(ham, nsec, ns)
    type(Total_Block), intent(inout) :: ham
    integer(c_int32_t), intent(in) :: nsec(:)
    integer, intent(in) :: ns
    type(Sub_Block), allocatable :: s(:)
    integer :: q, k
    if (ham%num /= ns) error stop 'Get_Hamiltonian: a new sector has no Hamiltonian block'
    allocate(s(ns))
    do q = 1, ns
      do k = 1, ham%num
        if (ham%sub(k)%num_up == nsec(3 * q - 2) .and. ham%sub(k)%num_down == nsec(3 * q - 1)) s(q) = ham%sub(k)
      end do
    end do
    call move_alloc(s, ham%sub)
  end subroutine sort_ham_by_sector

  ! the sys block of length L (sites 1..L) grown by site L+1 and truncated
  subroutine Get_Hamiltonian_Sys(pre, next, basis, trun)
    type(Total_Model), intent(inout) :: pre, next
    type(Total_Basis), intent(in) :: basis
    type(Total_Block), intent(in) :: trun
    call renormalize(pre, next, basis, trun, pre%len + 1, .false.)
  end subroutine Get_Hamiltonian_Sys

  ! the env block of length L (sites N..N-L+1) grown by site N-L and truncated
  subroutine Get_Hamiltonian_Env(pre, next, basis, trun)
    type(Total_Model), intent(inout) :: pre, next
    type(Total_Basis), intent(in) :: basis
    type(Total_Block), intent(in) :: trun
    call renormalize(pre, next, basis, trun, num_site - pre%len, .true.)
  end subroutine Get_Hamiltonian_Env

  ! identity truncation of an enlarged basis (every state kept)
  subroutine identity_trun(basis, trun)
    type(Total_Basis), intent(in) :: basis
    type(Total_Block), intent(inout) :: trun
    integer :: x, i
    call new_block(trun, 0, basis%num)
    do x = 1, basis%num
      trun%sub(x)%num_up = basis%sub(x)%new_num_up
      trun%sub(x)%num_down = basis%sub(x)%new_num_down
      trun%sub(x)%row_dim = basis%sub(x)%dim
      trun%sub(x)%sdim = basis%sub(x)%dim
      allocate(trun%sub(x)%mat(basis%sub(x)%dim, basis%sub(x)%dim))
      trun%sub(x)%mat = (0.0d0, 0.0d0)
      do i = 1, basis%sub(x)%dim
        trun%sub(x)%mat(i, i) = (1.0d0, 0.0d0)
      end do
    end do
  end subroutine identity_trun

end module tsc_dmrg
