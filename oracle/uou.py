"""U^H O U renormalisation (oracle for update_trun_dia / update_trun_ndia,
hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/block.f90:3465-3683): per operator
sub-block, mid = O . U_rhs (ZGEMM N,N) then eff = U_lhs^H . mid (ZGEMM C,N).
"""


def update_trun(blocks):
    return [Ul.conj().T @ (O @ Ur) for O, Ul, Ur in blocks]
