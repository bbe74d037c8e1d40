!----------------------------------------------------------------------------- best with 100 columns

!> ISO_C_BINDING interface to libfosolvers_amd.so (include/fosolvers_amd.h) for Fortran hosts such as
!> foeuler: the mesh build, condition and UDF tables, the Euler context (the CVODE callbacks of
!> src/foeuler/foeuler.f90:349-437), the integrator that replaces type(BDFNewtonKrylov)
!> (src/libfosolvers/numerics/SUNDIALS.f90:65-74,745-845) and the HIP runtime calls a host needs
!> for device vectors.  Every array behind a c_ptr is 0-based on the C side.
module modFosAmd
  use iso_c_binding
  implicit none
  public

  ! shape codes (polyGrid.f90:10-21), condition codes (foeuler.f90:17-25), context flags
  integer(c_int),parameter::FOS_TRI=10,FOS_QUAD=11,FOS_TET=20,FOS_HEX=21,FOS_TRI6=110,FOS_TET10=120
  integer(c_int),parameter::FOS_BC_WALL=0,FOS_BC_IN_STATIC=10,FOS_BC_IN_STATIC_UDF=110,&
  &                         FOS_BC_IN_TOTAL=11,FOS_BC_IN_TOTAL_UDF=111,FOS_BC_OUT=20,&
  &                         FOS_BC_OUT_UDF=120,FOS_BC_FAR=30,FOS_BC_FAR_UDF=130
  integer(c_int),parameter::FOS_VECTORS_ON_HOST=1,FOS_PREC_FP32=2,FOS_LAYOUT_GENERIC=4,FOS_JAC_EXACT=8,&
                          &FOS_ASYNC_CALLBACKS=16
  integer(c_int),parameter::FOS_RENUMBER_RCM=1,FOS_RENUMBER_CHUNKS=2
  ! hipMemcpyKind
  integer(c_int),parameter::HIP_H2D=1,HIP_D2H=2,HIP_D2D=3

  !> raw element lists (readGMSH's polyGrid)
  type,bind(c)::fos_poly_grid
    integer(c_int)::nN=0,nE=0,maxNNE=0
    type(c_ptr)::pN=c_null_ptr,sE=c_null_ptr,nNE=c_null_ptr,iNE=c_null_ptr,gid=c_null_ptr
  end type

  !> polyFvGrid after up(): cells [0,nC), facets [nC,nE); boundary pairs last
  type,bind(c)::fos_fv_grid
    integer(c_int)::nN=0,nE=0,nC=0,nF=0,nP=0,nPi=0,nPb=0,maxNF=0,maxNNE=0
    type(c_ptr)::pN=c_null_ptr,sE=c_null_ptr,nNE=c_null_ptr,iNE=c_null_ptr,gid=c_null_ptr
    type(c_ptr)::p=c_null_ptr,v=c_null_ptr,iEP=c_null_ptr,aP=c_null_ptr,normP=c_null_ptr
    type(c_ptr)::pP=c_null_ptr,neib=c_null_ptr
  end type

  !> condTab (condition.f90:8-16): p(maxPar,n) column-major
  type,bind(c)::fos_cond_tab
    integer(c_int)::n=0,maxPar=0
    type(c_ptr)::gid=c_null_ptr,t=c_null_ptr,p=c_null_ptr
  end type

  !> integrator counters
  type,bind(c)::fos_bdf_stats
    integer(c_long_long)::nst,nfe,nni,nli,npe,nps,njv,nncfails,nlinfails,nfallback
    integer(c_int)::lastNewtonIters,lastLinIters
    real(c_double)::lastResidualNorm
    integer(c_long_long)::nreorth,netf
    integer(c_int)::lastOrder,nextOrder
    real(c_double)::lastStep,nextStep
    real(c_double)::lastResidualNormCvode
  end type

  interface
    ! ---- mesh
    function fos_grid_read_gmsh(path,g) bind(c,name='fos_grid_read_gmsh')
      import; integer(c_int)::fos_grid_read_gmsh
      character(kind=c_char),intent(in)::path(*); type(fos_poly_grid)::g
    end function
    function fos_grid_write_gmsh(path,g) bind(c,name='fos_grid_write_gmsh')
      import; integer(c_int)::fos_grid_write_gmsh
      character(kind=c_char),intent(in)::path(*); type(fos_poly_grid),intent(in)::g
    end function
    function fos_grid_renumber(g,method,perm) bind(c,name='fos_grid_renumber')
      import; integer(c_int)::fos_grid_renumber
      type(fos_poly_grid)::g; integer(c_int),value::method; type(c_ptr),value::perm
    end function
    function fos_grid_up(g,fv) bind(c,name='fos_grid_up')
      import; integer(c_int)::fos_grid_up
      type(fos_poly_grid),intent(in)::g; type(fos_fv_grid)::fv
    end function
    function fos_grid_up_gpu(g,fv,stream) bind(c,name='fos_grid_up_gpu')
      import; integer(c_int)::fos_grid_up_gpu
      type(fos_poly_grid),intent(in)::g; type(fos_fv_grid)::fv; type(c_ptr),value::stream
    end function
    subroutine fos_poly_grid_free(g) bind(c,name='fos_poly_grid_free')
      import; type(fos_poly_grid)::g
    end subroutine
    subroutine fos_fv_grid_free(g) bind(c,name='fos_fv_grid_free')
      import; type(fos_fv_grid)::g
    end subroutine
    ! ---- conditions and UDFs (libmatheval's entry points, UDF.f90:24-46)
    function fos_read_cond_tab(path,c) bind(c,name='fos_read_cond_tab')
      import; integer(c_int)::fos_read_cond_tab
      character(kind=c_char),intent(in)::path(*); type(fos_cond_tab)::c
    end function
    subroutine fos_cond_tab_free(c) bind(c,name='fos_cond_tab_free')
      import; type(fos_cond_tab)::c
    end subroutine
    function fos_read_udf_tab(path,handles,maxHandles) bind(c,name='fos_read_udf_tab')
      import; integer(c_int)::fos_read_udf_tab
      character(kind=c_char),intent(in)::path(*); type(c_ptr)::handles(*); integer(c_int),value::maxHandles
    end function
    function evaluator_evaluate(f,n,names,vals) bind(c,name='evaluator_evaluate')
      import; real(c_double)::evaluator_evaluate
      type(c_ptr),value::f; integer(c_int),value::n; type(c_ptr)::names(*); real(c_double)::vals(*)
    end function
    subroutine evaluator_destroy(f) bind(c,name='evaluator_destroy')
      import; type(c_ptr),value::f
    end subroutine
    ! ---- Euler context and the CVODE callbacks (foeuler.f90:349-437)
    function fos_euler_create(grid,bc,r,gamm,flags,stream) bind(c,name='fos_euler_create')
      import; type(c_ptr)::fos_euler_create
      type(fos_fv_grid),intent(in)::grid; type(fos_cond_tab),intent(in)::bc
      real(c_double),value::r,gamm; integer(c_int),value::flags; type(c_ptr),value::stream
    end function
    subroutine fos_euler_destroy(ctx) bind(c,name='fos_euler_destroy')
      import; type(c_ptr),value::ctx
    end subroutine
    function fos_euler_set_tolerances(ctx,rtol,atol) bind(c,name='fos_euler_set_tolerances')
      import; integer(c_int)::fos_euler_set_tolerances
      type(c_ptr),value::ctx; real(c_double),value::rtol,atol
    end function
    function fos_euler_set_udf_bc(ctx,vals) bind(c,name='fos_euler_set_udf_bc')
      import; integer(c_int)::fos_euler_set_udf_bc
      type(c_ptr),value::ctx; real(c_double),intent(in)::vals(*)
    end function
    function fos_euler_sync_state(ctx,y,rho,rhou,rhoE,u,p,temp,c) bind(c,name='fos_euler_sync_state')
      import; integer(c_int)::fos_euler_sync_state
      type(c_ptr),value::ctx,y,rho,rhou,rhoE,u,p,temp,c
    end function
    function fos_euler_rhs(t,y,ydot,ud) bind(c,name='fos_euler_rhs')
      import; integer(c_int)::fos_euler_rhs
      real(c_double),value::t; type(c_ptr),value::y,ydot,ud
    end function
    function fos_euler_jtimes(v,jv,t,y,fy,ud,tmp) bind(c,name='fos_euler_jtimes')
      import; integer(c_int)::fos_euler_jtimes
      type(c_ptr),value::v,jv,y,fy,ud,tmp; real(c_double),value::t
    end function
    function fos_euler_jtimes_analytic(v,jv,t,y,fy,ud,tmp) bind(c,name='fos_euler_jtimes_analytic')
      import; integer(c_int)::fos_euler_jtimes_analytic
      type(c_ptr),value::v,jv,y,fy,ud,tmp; real(c_double),value::t
    end function
    !> out = v - gamma J v in one pass (no reference counterpart; for fos_bdf_set_newton_matvec)
    function fos_euler_newton_matvec(gamm,v,out,t,y,fy,ud) bind(c,name='fos_euler_newton_matvec')
      import; integer(c_int)::fos_euler_newton_matvec
      type(c_ptr),value::v,out,y,fy,ud; real(c_double),value::gamm,t
    end function
    function fos_euler_newton_matvec_analytic(gamm,v,out,t,y,fy,ud) bind(c,name='fos_euler_newton_matvec_analytic')
      import; integer(c_int)::fos_euler_newton_matvec_analytic
      type(c_ptr),value::v,out,y,fy,ud; real(c_double),value::gamm,t
    end function
    function fos_euler_psetup(t,y,fy,jok,jcur,gamm,ud) bind(c,name='fos_euler_psetup')
      import; integer(c_int)::fos_euler_psetup
      real(c_double),value::t,gamm; type(c_ptr),value::y,fy,ud; integer(c_int),value::jok
      integer(c_int)::jcur
    end function
    function fos_euler_psolve(t,y,fy,r,z,gamm,delta,lr,ud) bind(c,name='fos_euler_psolve')
      import; integer(c_int)::fos_euler_psolve
      real(c_double),value::t,gamm,delta; type(c_ptr),value::y,fy,r,z,ud; integer(c_int),value::lr
    end function
    ! ---- the same callbacks with SUNDIALS' signatures (N_Vector handles), for CVodeInit,
    !      CVodeSetJacTimes and CVodeSetPreconditioner; the host registers N_VGetArrayPointer once
    function fos_euler_set_nvector_access(ctx,getData) bind(c,name='fos_euler_set_nvector_access')
      import; integer(c_int)::fos_euler_set_nvector_access
      type(c_ptr),value::ctx; type(c_funptr),value::getData
    end function
    function fos_euler_rhs_nv(t,y,ydot,ud) bind(c,name='fos_euler_rhs_nv')
      import; integer(c_int)::fos_euler_rhs_nv
      real(c_double),value::t; type(c_ptr),value::y,ydot,ud
    end function
    function fos_euler_jtimes_nv(v,jv,t,y,fy,ud,tmp) bind(c,name='fos_euler_jtimes_nv')
      import; integer(c_int)::fos_euler_jtimes_nv
      type(c_ptr),value::v,jv,y,fy,ud,tmp; real(c_double),value::t
    end function
    function fos_euler_jtimes_analytic_nv(v,jv,t,y,fy,ud,tmp) bind(c,name='fos_euler_jtimes_analytic_nv')
      import; integer(c_int)::fos_euler_jtimes_analytic_nv
      type(c_ptr),value::v,jv,y,fy,ud,tmp; real(c_double),value::t
    end function
    function fos_euler_psetup_nv(t,y,fy,jok,jcur,gamm,ud) bind(c,name='fos_euler_psetup_nv')
      import; integer(c_int)::fos_euler_psetup_nv
      real(c_double),value::t,gamm; type(c_ptr),value::y,fy,ud; integer(c_int),value::jok; integer(c_int)::jcur
    end function
    function fos_euler_psolve_nv(t,y,fy,r,z,gamm,delta,lr,ud) bind(c,name='fos_euler_psolve_nv')
      import; integer(c_int)::fos_euler_psolve_nv
      real(c_double),value::t,gamm,delta; type(c_ptr),value::y,fy,r,z,ud; integer(c_int),value::lr
    end function
    function fos_euler_nonfinite(ctx,clear) bind(c,name='fos_euler_nonfinite')
      import; integer(c_int)::fos_euler_nonfinite
      type(c_ptr),value::ctx; integer(c_int),value::clear
    end function
    ! ---- writeState from the device state (foeuler.f90:333-346)
    function fos_euler_write_vtk(ctx,g,y,path) bind(c,name='fos_euler_write_vtk')
      import; integer(c_int)::fos_euler_write_vtk
      type(c_ptr),value::ctx; type(fos_fv_grid),intent(in)::g; type(c_ptr),value::y
      character(kind=c_char),intent(in)::path(*)
    end function
    ! ---- sortPolyFvGrid's exact pair order (polyFvGrid.f90:201-227)
    function fos_grid_sort_pairs_reference(fv) bind(c,name='fos_grid_sort_pairs_reference')
      import; integer(c_int)::fos_grid_sort_pairs_reference
      type(fos_fv_grid)::fv
    end function
    ! ---- the ILU type (sparse.f90:52-63): init, setCOO, setCSR, mulVec, fact, solve (0-based indices)
    function fos_ilut_create(nEq,maxNNz,lFill,flags,stream) bind(c,name='fos_ilut_create')
      import; type(c_ptr)::fos_ilut_create
      integer(c_int),value::nEq,maxNNz,lFill,flags; type(c_ptr),value::stream
    end function
    subroutine fos_ilut_destroy(p) bind(c,name='fos_ilut_destroy')
      import; type(c_ptr),value::p
    end subroutine
    function fos_ilut_set_coo(p,iA,jA,A,nNz) bind(c,name='fos_ilut_set_coo')
      import; integer(c_int)::fos_ilut_set_coo
      type(c_ptr),value::p; integer(c_int),intent(in)::iA(*),jA(*); real(c_double),intent(in)::A(*)
      integer(c_int),value::nNz
    end function
    function fos_ilut_set_csr(p,iA,jA,A) bind(c,name='fos_ilut_set_csr')
      import; integer(c_int)::fos_ilut_set_csr
      type(c_ptr),value::p; integer(c_int),intent(in)::iA(*),jA(*); real(c_double),intent(in)::A(*)
    end function
    function fos_ilut_mul_vec(p,x,y) bind(c,name='fos_ilut_mul_vec')
      import; integer(c_int)::fos_ilut_mul_vec
      type(c_ptr),value::p,x,y
    end function
    function fos_ilut_fact(p,droptol) bind(c,name='fos_ilut_fact')
      import; integer(c_int)::fos_ilut_fact
      type(c_ptr),value::p; real(c_double),value::droptol
    end function
    function fos_ilut_solve(p,rhs,x) bind(c,name='fos_ilut_solve')
      import; integer(c_int)::fos_ilut_solve
      type(c_ptr),value::p,rhs,x
    end function
    ! ---- integrator (BDFNewtonKrylov: init, setTol, setMaxSteps, setIV, step, getDt)
    function fos_bdf_create(nEq,f,maxl,pset,psol,ud) bind(c,name='fos_bdf_create')
      import; type(c_ptr)::fos_bdf_create
      integer(c_int),value::nEq,maxl; type(c_funptr),value::f,pset,psol; type(c_ptr),value::ud
    end function
    subroutine fos_bdf_destroy(b) bind(c,name='fos_bdf_destroy')
      import; type(c_ptr),value::b
    end subroutine
    function fos_bdf_set_jtimes(b,jt) bind(c,name='fos_bdf_set_jtimes')
      import; integer(c_int)::fos_bdf_set_jtimes; type(c_ptr),value::b; type(c_funptr),value::jt
    end function
    !> the Newton matrix product v - gamma J v in one pass (fos_euler_newton_matvec[_analytic])
    function fos_bdf_set_newton_matvec(b,fn) bind(c,name='fos_bdf_set_newton_matvec')
      import; integer(c_int)::fos_bdf_set_newton_matvec; type(c_ptr),value::b; type(c_funptr),value::fn
    end function
    function fos_bdf_set_options(b,flags,stream) bind(c,name='fos_bdf_set_options')
      import; integer(c_int)::fos_bdf_set_options; type(c_ptr),value::b; integer(c_int),value::flags
      type(c_ptr),value::stream
    end function
    function fos_bdf_set_tol(b,rtol,atol) bind(c,name='fos_bdf_set_tol')
      import; integer(c_int)::fos_bdf_set_tol; type(c_ptr),value::b; real(c_double),value::rtol,atol
    end function
    function fos_bdf_set_max_steps(b,n) bind(c,name='fos_bdf_set_max_steps')
      import; integer(c_int)::fos_bdf_set_max_steps; type(c_ptr),value::b; integer(c_int),value::n
    end function
    function fos_bdf_set_dt(b,dt) bind(c,name='fos_bdf_set_dt')
      import; integer(c_int)::fos_bdf_set_dt; type(c_ptr),value::b; real(c_double),value::dt
    end function
    function fos_bdf_set_adaptive(b,maxOrd,h0,hmin,hmax) bind(c,name='fos_bdf_set_adaptive')
      import; integer(c_int)::fos_bdf_set_adaptive; type(c_ptr),value::b
      integer(c_int),value::maxOrd; real(c_double),value::h0,hmin,hmax
    end function
    function fos_bdf_set_iv(b,t,y0) bind(c,name='fos_bdf_set_iv')
      import; integer(c_int)::fos_bdf_set_iv; type(c_ptr),value::b; real(c_double),value::t
      type(c_ptr),value::y0
    end function
    function fos_bdf_step(b,tOut,t,y) bind(c,name='fos_bdf_step')
      import; integer(c_int)::fos_bdf_step; type(c_ptr),value::b; real(c_double),value::tOut
      real(c_double)::t; type(c_ptr),value::y
    end function
    function fos_bdf_get_dt(b,dt) bind(c,name='fos_bdf_get_dt')
      import; integer(c_int)::fos_bdf_get_dt; type(c_ptr),value::b; real(c_double)::dt
    end function
    function fos_bdf_get_dky(b,t,k,dky) bind(c,name='fos_bdf_get_dky')
      import; integer(c_int)::fos_bdf_get_dky; type(c_ptr),value::b,dky; real(c_double),value::t
      integer(c_int),value::k
    end function
    function fos_bdf_get_stats(b,st) bind(c,name='fos_bdf_get_stats')
      import; integer(c_int)::fos_bdf_get_stats; type(c_ptr),value::b; type(fos_bdf_stats)::st
    end function
    function fos_last_error_c() bind(c,name='fos_last_error')
      import; type(c_ptr)::fos_last_error_c
    end function
    ! ---- HIP runtime (libamdhip64) for device vectors
    function hipMalloc(ptr,nbytes) bind(c,name='hipMalloc')
      import; integer(c_int)::hipMalloc; type(c_ptr)::ptr; integer(c_size_t),value::nbytes
    end function
    function hipFree(ptr) bind(c,name='hipFree')
      import; integer(c_int)::hipFree; type(c_ptr),value::ptr
    end function
    function hipMemcpy(dst,src,nbytes,kind) bind(c,name='hipMemcpy')
      import; integer(c_int)::hipMemcpy; type(c_ptr),value::dst,src
      integer(c_size_t),value::nbytes; integer(c_int),value::kind
    end function
    function hipDeviceSynchronize() bind(c,name='hipDeviceSynchronize')
      import; integer(c_int)::hipDeviceSynchronize
    end function
  end interface

contains

  !> the library's last error message as a Fortran string
  function fos_last_error() result(msg)
    character(len=:),allocatable::msg
    type(c_ptr)::p
    character(kind=c_char),pointer::s(:)
    integer::n
    p=fos_last_error_c()
    msg=''
    if(.not.c_associated(p))return
    call c_f_pointer(p,s,[4096])
    n=0
    do while(n<4096)
      if(s(n+1)==c_null_char)exit
      n=n+1
    end do
    allocate(character(len=n)::msg)
    do n=1,len(msg)
      msg(n:n)=s(n)
    end do
  end function

  !> a Fortran string as a NUL-terminated C string
  function c_str(s) result(r)
    character(*),intent(in)::s
    character(kind=c_char,len=len_trim(s)+1)::r
    r=trim(s)//c_null_char
  end function

  !> evaluate UDF handle h at position x and time t (UDF.f90:75-92: names x, y, z, t)
  function fos_udf_eval(h,x,t) result(v)
    type(c_ptr),intent(in)::h
    real(c_double),intent(in)::x(3),t
    real(c_double)::v
    character(len=2,kind=c_char),target,save::xName='x'//c_null_char,yName='y'//c_null_char,&
    &                                         zName='z'//c_null_char,tName='t'//c_null_char
    type(c_ptr)::names(4)
    real(c_double)::vals(4)
    names=[c_loc(xName),c_loc(yName),c_loc(zName),c_loc(tName)]
    vals=[x(1),x(2),x(3),t]
    v=evaluator_evaluate(h,4,names,vals)
  end function

end module
