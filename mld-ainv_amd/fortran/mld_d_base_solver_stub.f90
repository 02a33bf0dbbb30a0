! Minimal stand-in for MLD2P4's mld_d_base_solver_type so that the solver
! objects of mld_d_ainv_hip_mod build without MLD2P4.  Against a real
! MLD2P4 installation, mld_d_ainv_hip_mod uses mld_d_base_solver_mod
! instead and the types extend the framework's base solver directly.
module mld_d_base_solver_stub
  implicit none
  type :: mld_d_base_solver_type
  end type mld_d_base_solver_type
end module mld_d_base_solver_stub
