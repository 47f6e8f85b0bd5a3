#pragma once
// Field sets of the fused forward-mode passes (FISP EPG and pSSFP
// isochromats): a bit per tissue parameter whose tangent the pass carries,
// plus one for the echoes themselves.  Output planes of a pass are ordered
// echoes (when asked), then T1, T2, B1, B0 as present.

namespace compas {
namespace kernels {

enum : int { TM_T1 = 1, TM_T2 = 2, TM_B1 = 4, TM_B0 = 8, TM_ECHO = 16 };

template<int MASK>
struct TangentSet {
    static constexpr bool t1 = (MASK & TM_T1) != 0, t2 = (MASK & TM_T2) != 0;
    static constexpr bool b1 = (MASK & TM_B1) != 0, b0 = (MASK & TM_B0) != 0;
    static constexpr bool echo = (MASK & TM_ECHO) != 0;
    static constexpr int nt = int(t1) + int(t2) + int(b1) + int(b0);  // tangents carried
    static constexpr int planes = nt + int(echo);
    static constexpr int it1 = 0, it2 = int(t1), ib1 = int(t1) + int(t2), ib0 = int(t1) + int(t2) + int(b1);
    static constexpr bool axis = b1 || b0;  // tangents that move the pSSFP rotation axis
};

}  // namespace kernels
}  // namespace compas
