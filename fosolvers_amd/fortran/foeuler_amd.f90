!----------------------------------------------------------------------------- best with 100 columns

!> foeuler on libfosolvers_amd.  Reads grid.msh, bc, ic, sim, fl (and udf when a UDF is used)
!> from the working directory with foeuler's formats and meaning (src/foeuler/foeuler.f90:60-175),
!> builds the finite-volume grid on the GPU (the same arrays as polyFvGrid%up), integrates the
!> Euler equations with the library's device-resident BDF-Newton-GMRES and block ILU(0) from t = 0
!> to tFinal one internal step at a time (:442-463), and writes rst_N.vtk at
!> every output interval (writeState, :333-346).  RTOL/ATOL are foeuler's (:14-15).
!>
!> Environment: FOS_RENUMBER=1 renumbers the cells (reverse Cuthill-McKee, then 256-cell chunks)
!> before the grid is built, for memory locality on large meshes; the VTK files then list the
!> cells in the new order.  FOS_FP32=1 computes the flux, Jv and preconditioner apply in FP32
!> (Newton residual and update stay FP64), FOS_JV_ANALYTIC=1 replaces CVODE's difference-quotient
!> Jv by the exact product, FOS_JAC_EXACT=1 factors the exact Jacobian instead of findEulerJac's,
!> FOS_LAYOUT_GENERIC=1 evaluates the residual with the generic face layout, whose per-cell flux
!> sums run in a fixed order (bitwise reproducible runs; the tetrahedral layout's LDS atomic sums
!> agree to rounding).
program foeuler_amd
  use modFosAmd
  implicit none
  double precision,parameter::RTOL=1d-6,ATOL=1d-12
  integer,parameter::FID=10,MAX_UDF=4096
  type(fos_poly_grid)::pg
  type(fos_fv_grid)::fv
  type(fos_cond_tab)::bc
  type(fos_bdf_stats)::st
  type(c_ptr)::ctx,ode,yDev,eDev(7)
  type(c_ptr)::udf(MAX_UDF)
  integer::nUdf,udfIc,iUdf(5),iOut,i,j,k,nC,nE,nEq,nPi,nPb,rc
  logical::udfBc
  double precision::p0,T0,u0(3),pE(3),tFinal,tInt,r,gamm,t,tNext,dt,tmpD
  double precision,allocatable,target::rho(:),rhou(:,:),rhoE(:),u(:,:),p(:),temp(:),c(:),y(:)
  double precision,allocatable::udfVals(:)
  integer,allocatable::iBC(:)
  real(c_double),pointer::pN(:,:),pC(:,:),pP(:,:),bcP(:,:)
  integer(c_int),pointer::iEP(:,:),gid(:),sE(:),nNE(:),iNE(:,:),bcGid(:),bcT(:)
  character(20)::tmpStr
  character(8)::envVal
  integer(c_int)::ctxFlags

  nUdf=0
  ! ---- grid (foeuler.f90:67-70)
  rc=fos_grid_read_gmsh(c_str('grid.msh'),pg)
  if(rc/=0)call fail('reading grid.msh')
  call get_environment_variable('FOS_RENUMBER',envVal)
  if(trim(envVal)=='1')then
    if(fos_grid_renumber(pg,FOS_RENUMBER_RCM,c_null_ptr)/=0)call fail('renumbering (RCM)')
    if(fos_grid_renumber(pg,FOS_RENUMBER_CHUNKS,c_null_ptr)/=0)call fail('renumbering (chunks)')
  end if
  if(fos_grid_up_gpu(pg,fv,c_null_ptr)/=0)call fail('building the finite-volume grid')
  nC=fv%nC
  nE=fv%nE
  nPi=fv%nPi
  nPb=fv%nPb
  nEq=5*nC
  call c_f_pointer(fv%pN,pN,[3,fv%nN])
  call c_f_pointer(fv%p,pC,[3,nE])
  call c_f_pointer(fv%pP,pP,[3,fv%nP])
  call c_f_pointer(fv%iEP,iEP,[2,fv%nP])
  call c_f_pointer(fv%gid,gid,[nE])
  call c_f_pointer(fv%sE,sE,[nE])
  call c_f_pointer(fv%nNE,nNE,[nE])
  call c_f_pointer(fv%iNE,iNE,[fv%maxNNE,nE])
  ! ---- conditions and parameters (:71-112)
  if(fos_read_cond_tab(c_str('bc'),bc)/=0)call fail('reading bc')
  call c_f_pointer(bc%gid,bcGid,[bc%n])
  call c_f_pointer(bc%t,bcT,[bc%n])
  call c_f_pointer(bc%p,bcP,[bc%maxPar,bc%n])
  udfBc=any(bcT==FOS_BC_IN_STATIC_UDF.or.bcT==FOS_BC_IN_TOTAL_UDF.or.bcT==FOS_BC_OUT_UDF.or.&
  &         bcT==FOS_BC_FAR_UDF)
  open(FID,file='ic',action='read')
  read(FID,*)udfIc
  if(udfIc==0)then
    read(FID,*)p0
    read(FID,*)T0
    read(FID,*)u0(1)
    read(FID,*)u0(2)
    read(FID,*)u0(3)
  else
    do k=1,5
      read(FID,*)tmpD
      iUdf(k)=int(tmpD)
    end do
  end if
  close(FID)
  open(FID,file='sim',action='read')
  read(FID,*)tFinal
  read(FID,*)tInt
  close(FID)
  open(FID,file='fl',action='read')
  read(FID,*)r
  read(FID,*)gamm
  close(FID)
  if(udfIc/=0.or.udfBc)then
    nUdf=fos_read_udf_tab(c_str('udf'),udf,MAX_UDF)
    if(nUdf<0)call fail('reading udf')
  end if
  ! ---- initial state on every element, cells and facets (:113-133)
  allocate(rho(nE),rhou(3,nE),rhoE(nE),u(3,nE),p(nE),temp(nE),c(nE))
  do i=1,nE
    if(udfIc/=0)then
      pE(:)=pC(:,i)
      p0=eval_udf(iUdf(1),pE,0d0)
      T0=eval_udf(iUdf(2),pE,0d0)
      u0(1)=eval_udf(iUdf(3),pE,0d0)
      u0(2)=eval_udf(iUdf(4),pE,0d0)
      u0(3)=eval_udf(iUdf(5),pE,0d0)
    end if
    rho(i)=p0/r/T0
    rhou(:,i)=rho(i)*u0(:)
    rhoE(i)=rho(i)*(1d0/(gamm-1d0)*r*T0+0.5d0*dot_product(u0,u0))
    u(:,i)=u0(:)
    p(i)=p0
    temp(i)=T0
    c(i)=sqrt(gamm*r*T0)
  end do
  t=0d0
  tNext=tInt
  iOut=0
  ! condition index of every element: first matching gid, 0 = none (wall) (:137-147)
  allocate(iBC(nE))
  iBC(:)=0
  do i=1,nE
    do j=1,bc%n
      if(gid(i)==bcGid(j))then
        iBC(i)=j
        exit
      end if
    end do
  end do
  ! ---- state vector, context, integrator (:148-174)
  allocate(y(nEq))
  y(1:nC)=rho(1:nC)
  y(nC+1:2*nC)=rhou(1,1:nC)
  y(2*nC+1:3*nC)=rhou(2,1:nC)
  y(3*nC+1:4*nC)=rhou(3,1:nC)
  y(4*nC+1:5*nC)=rhoE(1:nC)
  ctxFlags=0
  if(env_on('FOS_FP32'))ctxFlags=ior(ctxFlags,FOS_PREC_FP32)
  if(env_on('FOS_JAC_EXACT'))ctxFlags=ior(ctxFlags,FOS_JAC_EXACT)
  if(env_on('FOS_LAYOUT_GENERIC'))ctxFlags=ior(ctxFlags,FOS_LAYOUT_GENERIC)
  ctx=fos_euler_create(fv,bc,r,gamm,ctxFlags,c_null_ptr)
  if(.not.c_associated(ctx))call fail('creating the Euler context')
  if(fos_euler_set_tolerances(ctx,RTOL,ATOL)/=0)call fail('setting tolerances')
  if(udfBc)then
    allocate(udfVals(5*max(nPb,1)))
    call set_udf_bc()
  end if
  call dev_alloc(yDev,nEq)
  call h2d(yDev,c_loc(y),nEq)
  call dev_alloc(eDev(1),nE)
  call dev_alloc(eDev(2),3*nE)
  call dev_alloc(eDev(3),nE)
  call dev_alloc(eDev(4),3*nE)
  call dev_alloc(eDev(5),nE)
  call dev_alloc(eDev(6),nE)
  call dev_alloc(eDev(7),nE)
  ! facets without a boundary pair keep their initial values, as in the reference
  call h2d(eDev(1),c_loc(rho),nE)
  call h2d(eDev(2),c_loc(rhou),3*nE)
  call h2d(eDev(3),c_loc(rhoE),nE)
  call h2d(eDev(4),c_loc(u),3*nE)
  call h2d(eDev(5),c_loc(p),nE)
  call h2d(eDev(6),c_loc(temp),nE)
  call h2d(eDev(7),c_loc(c),nE)
  ode=fos_bdf_create(nEq,c_funloc(fos_euler_rhs),0,c_funloc(fos_euler_psetup),&
  &                  c_funloc(fos_euler_psolve),ctx)
  if(.not.c_associated(ode))call fail('creating the integrator')
  ! the Jv of GMRES products fused with the Newton combine v - gamma J v (one pass)
  if(env_on('FOS_JV_ANALYTIC'))then
    if(fos_bdf_set_jtimes(ode,c_funloc(fos_euler_jtimes_analytic))/=0)call fail('setting jtimes')
    if(fos_bdf_set_newton_matvec(ode,c_funloc(fos_euler_newton_matvec_analytic))/=0)call fail('setting jtimes')
  else
    if(fos_bdf_set_jtimes(ode,c_funloc(fos_euler_jtimes))/=0)call fail('setting jtimes')
    if(fos_bdf_set_newton_matvec(ode,c_funloc(fos_euler_newton_matvec))/=0)call fail('setting jtimes')
  end if
  ! mixed precision: the GMRES basis in FP32 too
  if(iand(ctxFlags,FOS_PREC_FP32)/=0)then
    if(fos_bdf_set_options(ode,FOS_PREC_FP32,c_null_ptr)/=0)call fail('setting integrator options')
  end if
  if(fos_bdf_set_tol(ode,RTOL,ATOL)/=0)call fail('setting tolerances')
  if(fos_bdf_set_max_steps(ode,-1)/=0)call fail('setting max steps')
  if(fos_bdf_set_iv(ode,t,yDev)/=0)call fail('setting the initial value')
  ! ---- time loop (:442-463)
  write(tmpStr,*)iOut
  call writeState('rst_'//trim(adjustl(tmpStr))//'.vtk')
  do while(t<tFinal)
    rc=fos_bdf_step(ode,tNext,t,yDev)
    if(rc/=0)call fail('time step')
    rc=fos_bdf_get_dt(ode,dt)
    write(*,'(a,g12.5,a,g12.5)')'[i] finished time step: t=',t,' dt=',dt
    if(udfBc)call set_udf_bc()
    if(t+tiny(1d0)>=tNext)then
      iOut=iOut+1
      write(tmpStr,*)iOut
      call syncState()
      write(*,'(a)')'[i] writing: rst_'//trim(adjustl(tmpStr))//'.vtk'
      call writeState('rst_'//trim(adjustl(tmpStr))//'.vtk')
      tNext=tNext+tInt
    end if
  end do
  rc=fos_bdf_get_stats(ode,st)
  write(*,'(a,i0,a,i0,a,i0,a,i0,a,i0,a,i0)')'[i] steps ',st%nst,', rhs calls ',st%nfe,&
  &  ', Newton iterations ',st%nni,', GMRES iterations ',st%nli,', preconditioner setups ',st%npe,&
  &  ', error test failures ',st%netf
  ! ---- clear (:177-183)
  call fos_bdf_destroy(ode)
  call fos_euler_destroy(ctx)
  rc=hipFree(yDev)
  do k=1,7
    rc=hipFree(eDev(k))
  end do
  do k=1,nUdf
    call evaluator_destroy(udf(k))
  end do
  call fos_cond_tab_free(bc)
  call fos_fv_grid_free(fv)
  call fos_poly_grid_free(pg)

contains

  !> an environment switch set to 1
  logical function env_on(name)
    character(*),intent(in)::name
    character(8)::val
    call get_environment_variable(name,val)
    env_on=trim(val)=='1'
  end function

  !> the k-th UDF (1-based) at x and time tt
  function eval_udf(kk,x,tt) result(v)
    integer,intent(in)::kk
    double precision,intent(in)::x(3),tt
    double precision::v
    if(kk<1.or.kk>nUdf)call fail('UDF index out of range')
    v=fos_udf_eval(udf(kk),x,tt)
  end function

  !> UDF boundary parameters at the face centres and the time of the last accepted step, as
  !> setBC evaluates them (foeuler.f90:236-299); 5 values per boundary pair in pair order
  subroutine set_udf_bc()
    integer::kk,ii,nn,jj,q,nq
    udfVals(:)=0d0
    do kk=1,nPb
      ii=nPi+kk
      nn=iEP(2,ii)+1
      jj=iBC(nn)
      if(jj==0)cycle
      select case(bcT(jj))
      case(FOS_BC_IN_STATIC_UDF,FOS_BC_IN_TOTAL_UDF,FOS_BC_FAR_UDF)
        nq=5
      case(FOS_BC_OUT_UDF)
        nq=1
      case default
        nq=0
      end select
      do q=1,min(nq,bc%maxPar)
        udfVals(5*(kk-1)+q)=eval_udf(int(bcP(q,jj)),pP(:,ii),t)
      end do
    end do
    if(fos_euler_set_udf_bc(ctx,udfVals)/=0)call fail('setting UDF boundary values')
  end subroutine

  !> setBC + syncState of the current state into the element arrays (device, then host)
  subroutine syncState()
    if(fos_euler_sync_state(ctx,yDev,eDev(1),eDev(2),eDev(3),eDev(4),eDev(5),eDev(6),eDev(7))/=0)&
    &  call fail('element state')
    call d2h(c_loc(rho),eDev(1),nE)
    call d2h(c_loc(rhou),eDev(2),3*nE)
    call d2h(c_loc(rhoE),eDev(3),nE)
    call d2h(c_loc(u),eDev(4),3*nE)
    call d2h(c_loc(p),eDev(5),nE)
    call d2h(c_loc(temp),eDev(6),nE)
    call d2h(c_loc(c),eDev(7),nE)
  end subroutine

  !> legacy VTK: the grid's nodes and elements, then density, velocity, pressure and temperature
  !> per element (writeVTK, src/libfosolvers/fileIO/VTK.f90:83-204)
  subroutine writeState(fName)
    character(*),intent(in)::fName
    integer::ii,vt
    open(FID,file=trim(fName),action='write')
    write(FID,'(a)')'# vtk DataFile Version 3.0'
    write(FID,'(a)')'Generated by fosolvers_amd'
    write(FID,'(a)')'ASCII'
    write(FID,'(a)')''
    write(FID,'(a)')'DATASET UNSTRUCTURED_GRID'
    write(FID,'(a,i0,a)')'POINTS ',fv%nN,' double'
    do ii=1,fv%nN
      write(FID,'(3es25.16e3)')pN(:,ii)
    end do
    write(FID,'(a)')''
    write(FID,'(a,i0,1x,i0)')'CELLS ',nE,nE+sum(nNE)
    do ii=1,nE
      if(sE(ii)==FOS_TET10)then  ! VTK's quadratic tetrahedron swaps the last two edge nodes
        write(FID,'(*(i0,:,1x))')nNE(ii),iNE([1,2,3,4,5,6,7,8,10,9],ii)
      else
        write(FID,'(*(i0,:,1x))')nNE(ii),iNE(1:nNE(ii),ii)
      end if
    end do
    write(FID,'(a)')''
    write(FID,'(a,i0)')'CELL_TYPES ',nE
    do ii=1,nE
      select case(sE(ii))
      case(FOS_TRI)
        vt=5
      case(FOS_QUAD)
        vt=9
      case(FOS_TRI6)
        vt=22
      case(FOS_TET)
        vt=10
      case(FOS_HEX)
        vt=12
      case(FOS_TET10)
        vt=24
      case default
        vt=0
      end select
      write(FID,'(i0)')vt
    end do
    write(FID,'(a)')''
    write(FID,'(a,i0)')'CELL_DATA ',nE
    call scal('density',rho)
    call vect('velocity',u)
    call scal('pressure',p)
    call scal('temperature',temp)
    close(FID)
  end subroutine

  subroutine scal(key,a)
    character(*),intent(in)::key
    double precision,intent(in)::a(:)
    integer::ii
    write(FID,'(a)')''
    write(FID,'(a)')'SCALARS '//key//' double'
    write(FID,'(a)')'LOOKUP_TABLE default'
    do ii=1,size(a)
      write(FID,'(es25.16e3)')a(ii)
    end do
  end subroutine

  subroutine vect(key,a)
    character(*),intent(in)::key
    double precision,intent(in)::a(:,:)
    integer::ii
    write(FID,'(a)')''
    write(FID,'(a)')'VECTORS '//key//' double'
    do ii=1,size(a,2)
      write(FID,'(3es25.16e3)')a(:,ii)
    end do
  end subroutine

  subroutine dev_alloc(ptr,n)
    type(c_ptr),intent(out)::ptr
    integer,intent(in)::n
    if(hipMalloc(ptr,int(8,c_size_t)*int(max(n,1),c_size_t))/=0)call fail('device allocation')
  end subroutine

  subroutine h2d(dst,src,n)
    type(c_ptr),intent(in)::dst,src
    integer,intent(in)::n
    if(n>0)then
      if(hipMemcpy(dst,src,int(8,c_size_t)*int(n,c_size_t),HIP_H2D)/=0)call fail('copy to device')
    end if
  end subroutine

  subroutine d2h(dst,src,n)
    type(c_ptr),intent(in)::dst,src
    integer,intent(in)::n
    if(n>0)then
      if(hipMemcpy(dst,src,int(8,c_size_t)*int(n,c_size_t),HIP_D2H)/=0)call fail('copy to host')
    end if
  end subroutine

  !> the reference prints [E] and stops (SUNDIALS.f90:761-793); so does this driver, with the
  !> library's message
  subroutine fail(what)
    character(*),intent(in)::what
    write(*,'(a)')'[E] '//what//': '//fos_last_error()
    error stop 1
  end subroutine

end program
