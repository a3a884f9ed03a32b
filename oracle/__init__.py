"""CPU oracle for the DMRG hot path -- test infrastructure only.

A restatement, in numpy and plain Python, of the reference's algorithms
(hyx15903/TSC_Tri_t1t2J1J2, Fortran, DMRG_code/). Only tests/, the smoke
check of __graft_entry__ and the cpu_baseline leg of bench.py use it, and
only as the checker. The product (tsc_tri_t1t2j1j2_amd) never imports it.

Pinning (the reference itself cannot be built here: it needs Intel MKL,
which this image lacks, plus source edits to its compile-time lattice
parameters):
  wigner.w6j          exact rational Racah sum; checked against sympy's
                      wigner_6j, closed forms and Regge symmetries
  truncation.*        numpy eigh; invariants (Hermitian rho, trace 1,
                      orthonormal U, discarded weight = 1 - kept weight),
                      brute-force selection on tiny spectra
  lanczos.*           exact ground energies of dense Hermitian matrices
                      (numpy eigh)
  uou.update_trun     identity U maps O to O; equals U^H O U elementwise
  synth.hpsi          direct per-term product definition (no reordering)
  su2 / tj_model      exact diagonalisation of small t-J clusters in the
                      U(1) Fock basis (see tests/test_oracle_physics.py)
  tj_ed correlators   bond sums of <hopping>, <S.S>, <n n> rebuild the ED
                      energy; sum of <n_i> = N_e; sum of <S_i.S_j> = 0 for a
                      singlet; pair correlators equal overlaps of annihilated
                      states; the dimer singlet gives <D+ D> = 1
  ref_measure.*       the reference's coupling rule and file factors as
                      Fock-space matrices: on-site identities (S.S = 3/4 n,
                      n n = n, sum_s c+_s c_s = n), two-site operators equal
                      the physical ones; the standard Clebsch-Gordan phase in
                      its place flips the pair operator (the check fails)
  ref_blocks.*        the reference's block states as Fock vectors: every
                      enlarged state is a multiplet (each operator's
                      components fit the Wigner-Eckart form to 1e-10), the
                      block Hamiltonian's reduced matrix is that of the sum
                      of its bonds, the stored ground state is an exact
                      eigenvector of the whole-lattice H and stops being one
                      when the env sign is dropped; inb1 against the closed
                      rule (bonds leaving the block, plus sites whose last
                      neighbour is the block's last site); the restated
                      model_wave site-site hopping and Get_hamiltonian_ndia
                      block-site hopping are both 1/sqrt2 times the physical
                      hopping in these states (the opposite for the cross
                      term under the other superblock coupling order)
  ref_kernels.*       the 14 model_wave kernels restated from the Fortran:
                      each is one constant times its physical term on every
                      basis wave function of a 6-site superblock (residual
                      1e-12), the constant the couplings jt, jd, jn assume;
                      a dropped block_transfer_trans or env sign breaks it;
                      wave_trans_LR/RL keep a complex random state exactly
                      under complete complex unitary truncations, and fail
                      when env block states are read as unconjugated columns;
                      the six enlargement routines give the physical reduced
                      matrices in the enlarged basis on both sides, and the
                      block-site bond terms of Get_Hamiltonian_Sys/Env sum
                      to the physical bond
"""
