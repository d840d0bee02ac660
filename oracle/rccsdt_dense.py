"""Full-tensor restatement of the triples terms (oracle, test infrastructure only).

The block loop of ``rccsdt_tri`` reaches every term through case analysis over
canonical source/target triples (``distr_cc/rccsdt.py:253-732``).  Here the
same mathematics is written once over the dense ordered-triple tensor
T[i,j,k,a,b,c] (built from the canonical blocks with the simultaneous
permutation symmetry), with the case analysis resolved by hand:

* W_oooo/F_oo (``rccsdt.py:632-723`` comments):
      + W_oooo[lmij] T[lmk,abc] + W_oooo[lmik] T[lmj,acb] + W_oooo[lmjk] T[lmi,bca]
      - F_oo[li] T[ljk,abc] - F_oo[lj] T[ilk,abc] - F_oo[lk] T[ijl,abc]
* ring terms (``rccsdt.py:540-596``), derived from case 0 of part 3 by
  rewriting the canonical source block T[jkl] as T[ljk] with permuted
  virtuals (T[jkl][x,y,z] = T[ljk][z,x,y]):
      G[ijk,abc] = - W_ovov[lbid] T[ljk,adc] - W_ovov[lcid] T[ljk,abd]
                   - 1/2 W_ovov[laid] (T[ljk,bdc] + T[ljk,cbd])
                   + 1/2 W_ovvo[ladi] (2 T[ljk,dbc] - T[ljk,cbd] - T[ljk,bdc])
  parts 4 and 5 are G with (i,a)<->(j,b) and (i,a)<->(k,c):
      ring = G + G[jik,bac] + G[kji,cba]
* T3 -> r1/r2/W (``rccsdt.py:253-366``): sums over all ordered triples with
  Y = P3_422(T), Z = P3_201(T) applied per ordered block.

Agreement of this module with ``rccsdt_tri`` on random inputs checks the
reference's case analysis (cases 0/1/2, tie triples, batch boundaries) which
this module does not contain.
"""
import numpy as np

from oracle.spin_sum import t3_spin_sum


def _e(s, *ops):
    return np.einsum(s, *ops, optimize=True)


def r3_dense(imds, t2, T):
    """Dense r3 (before projection) for the full (o,o,o,v,v,v) tensor T."""
    F_oo, F_vv = imds['F_oo'], imds['F_vv']
    W_oooo, W_ovvo, W_ovov = imds['W_oooo'], imds['W_ovvo'], imds['W_ovov']
    W_vooo, W_vvvo, W_vvvv = imds['W_vooo'], imds['W_vvvo'], imds['W_vvvv']
    R = _e('abdj,ikdc->ijkabc', W_vvvo, t2)
    R += _e('acdk,ijdb->ijkabc', W_vvvo, t2)
    R += _e('badi,jkdc->ijkabc', W_vvvo, t2)
    R += _e('bcdk,jida->ijkabc', W_vvvo, t2)
    R += _e('cadi,kjdb->ijkabc', W_vvvo, t2)
    R += _e('cbdj,kida->ijkabc', W_vvvo, t2)
    R -= _e('alij,lkbc->ijkabc', W_vooo, t2)
    R -= _e('alik,ljcb->ijkabc', W_vooo, t2)
    R -= _e('blji,lkac->ijkabc', W_vooo, t2)
    R -= _e('bljk,lica->ijkabc', W_vooo, t2)
    R -= _e('clki,ljab->ijkabc', W_vooo, t2)
    R -= _e('clkj,liba->ijkabc', W_vooo, t2)
    R += _e('ad,ijkdbc->ijkabc', F_vv, T)
    R += _e('bd,ijkadc->ijkabc', F_vv, T)
    R += _e('cd,ijkabd->ijkabc', F_vv, T)
    R += _e('abde,ijkdec->ijkabc', W_vvvv, T)
    R += _e('acde,ijkdbe->ijkabc', W_vvvv, T)
    R += _e('bcde,ijkade->ijkabc', W_vvvv, T)
    R += _e('lmij,lmkabc->ijkabc', W_oooo, T)
    R += _e('lmik,lmjacb->ijkabc', W_oooo, T)
    R += _e('lmjk,lmibca->ijkabc', W_oooo, T)
    R -= _e('li,ljkabc->ijkabc', F_oo, T)
    R -= _e('lj,ilkabc->ijkabc', F_oo, T)
    R -= _e('lk,ijlabc->ijkabc', F_oo, T)
    G = -_e('lbid,ljkadc->ijkabc', W_ovov, T)
    G -= _e('lcid,ljkabd->ijkabc', W_ovov, T)
    G -= 0.5 * _e('laid,ljkbdc->ijkabc', W_ovov, T)
    G -= 0.5 * _e('laid,ljkcbd->ijkabc', W_ovov, T)
    G += 0.5 * _e('ladi,ljkdbc->ijkabc', W_ovvo, 2.0 * T)
    G -= 0.5 * _e('ladi,ljkcbd->ijkabc', W_ovvo, T)
    G -= 0.5 * _e('ladi,ljkbdc->ijkabc', W_ovvo, T)
    R += G + G.transpose(1, 0, 2, 4, 3, 5) + G.transpose(2, 1, 0, 5, 4, 3)
    return R


def r1r2_t3_dense(nocc, t1_fock, t1_eris, T):
    """T3 contributions to r1 and the un-symmetrized r2 (rccsdt.py:253-321)."""
    o = nocc
    Y = t3_spin_sum(T, "P3_422")
    Z = t3_spin_sum(T, "P3_201")
    r1 = 0.5 * _e('jkbc,ijkabc->ia', t1_eris[:o, :o, o:, o:], Y)
    r2 = 0.5 * _e('pa,pqrabc->qrbc', t1_fock[:o, o:], Z)
    r2 += _e('cpad,pqrdba->qrbc', t1_eris[o:, :o, o:, o:], Z)
    r2 -= _e('rpla,pqrabc->qlbc', t1_eris[:o, :o, :o, o:], Z)
    return r1, r2


def w_t3_dense(nocc, t1_eris, T):
    """T3 contributions to W_vooo and W_vvvo (rccsdt.py:323-366)."""
    o = nocc
    Z = t3_spin_sum(T, "P3_201")
    W_vooo = _e('mpde,pqread->amqr', t1_eris[:o, :o, o:, o:], Z)
    W_vvvo = -_e('rpde,pqreba->abdq', t1_eris[:o, :o, o:, o:], Z)
    return W_vooo, W_vvvo
